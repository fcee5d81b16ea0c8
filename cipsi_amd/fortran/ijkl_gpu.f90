! ijkl_gpu -- drop-in host program for the ijkl step, running the AO->MO transform on
! an MI355X through the C ABI of libcipsi_ijkl.so.  C1, the point groups without
! degenerate irreps (CS, CI, CN, CNV, CNH, DN, DNH with a two-fold axis) and the linear
! groups CINFV and DINFH with their symmetry transformations.
!
! It keeps ijkl's file contract so that fock and the rest of the chain run unchanged:
!   stdin   &ijfil (file names, CIPSI_sources/ijkl/openf.f:6-34) then &ijklin
!           (the same keys as ijkl.f:28-33, plus precision and nslice for this host)
!   unit 2  info:  record 1 header + basis and symmetry tables (read as rondof.f:19-38),
!                  record 2 S, record 3 h packed, record 7 C square (rotat2.f:33-38,
!                  tmono.f:19-22)
!   unit 8  pqrs:  NREC records of (8192 REAL*8, 8192 INT*4) labels (itijkl.f:57-68)
!   unit 10 ijom:  C twice and the S record (ijkl.f:223-240)
!   unit 25 f25:   header, its, block directory, enuc/ljkf (ijkl.f:336-340,406;
!                  itijkl.f:294-296)
!   unit 50 f50:   packed MO integrals in 8192-double records, the last written full
!                  length (tijkld.f:68-88), then the h_MO record (tmono.f:56)
! With symmetry the MO preparation of otodon / rotat2 / vector runs in the library
! (cipsi_ijkl_sym_prepare; cipsi_ijkl_sym_linear and cipsi_ijkl_sym_prepare_so for the
! linear groups) and the f50 stream holds the symmetry-allowed blocks in the order of
! itijkl.f:70-247, less those the symmetry transformations make redundant, with their
! directory in f25.  The unit-40 scratch file of
! the reference is not needed: the intermediate stays in device memory.  Errors are
! reported like ijkl's (' !! ERROR in ijkl', a message and an i2 tag read by
! Autocip_13/calcul.f90:169-176): rotat2's own tags (2, 1, -1) where it would stop,
! tag 9 for library failures and unsupported input.  istop = 1 stops after ijom, as
! ijkl.f:328-329 does.
program ijkl_gpu
  use iso_c_binding
  use cipsi_ijkl_mod
  implicit none
  integer, parameter :: lsize = 8192, doa = 512, nsymz = 20

  ! ---- &ijfil (openf.f:6-34)
  character(len=80) :: prefix
  character(len=10) :: f25, f50, info, pqrs, ijom, om, scr
  character(len=90) :: pqrs1, info1, ijom1, om1, f251, f501, scr1
  namelist /ijfil/ prefix, f25, f50, info, pqrs, ijom, om, scr, pqrs1, info1, ijom1, om1, &
                   f251, f501, scr1

  ! ---- &ijklin: every key of ijkl.f:28-33 is accepted; only the C1 subset acts
  integer :: nsym, norb, itsym(doa), itsyv(2*doa), isymp(nsymz), nvec, noc, ntrans, istop, &
             ljkf, nsyma, nsymb, nrec, mblk, mijkl, naxis, nveci, nvecf, ivecsy(2*doa*nsymz), &
             iwant, nocsy(nsymz), inpv
  logical(1) :: yhondo, yprtv, yauto, yrot, yuhf, ymono, yprt, yatom, ycv, ywvect, yprth, &
                ynocsy, yorbf, yrecsp, ygel(2*doa), yf11, ydp, yreduc
  logical :: chstan
  character(len=8) :: group
  double precision :: trec, tdeg
  ! keys of this host: precision 0 = FP64 MFMA, 1 = split FP16; nslice for the split path
  integer :: precision, nslice
  namelist /ijklin/ nsym, norb, itsym, itsyv, isymp, nvec, yhondo, noc, ntrans, yprtv, istop, &
                    group, yauto, ljkf, nsyma, nsymb, nrec, mblk, mijkl, naxis, yrot, yuhf, &
                    ymono, yprt, nveci, nvecf, ivecsy, yatom, iwant, ycv, ywvect, yprth, &
                    ynocsy, nocsy, yorbf, inpv, trec, tdeg, yrecsp, ygel, yf11, ydp, chstan, &
                    yreduc, precision, nslice

  ! ---- info header (ijkl.f:100-101) and basis tables (rondof.f:98-106)
  integer :: ioda(19), nt, nshell, num, nveca, nvecb, ngauss, ne, ich, mul, na, nb, nat, naxh
  character(len=8) :: bname, grouph
  double precision :: title(10), enuc
  integer, allocatable :: ibas(:), newsh(:), iso(:)
  integer :: invt(48)
  double precision, allocatable :: gauss(:), ptr(:), dtr(:), ftr(:)
  double precision, allocatable :: atom_c(:), atom_a(:), atom_zan(:)
  integer, allocatable :: atom_ns(:), atom_ks(:), atom_iat(:)
  character(len=4), allocatable :: iflab(:)
  character(len=6) :: ante

  double precision, allocatable :: c(:, :), h(:), s(:), xx(:), mo(:), hmo(:), buf(:)
  integer(c_int32_t), allocatable :: ix(:)
  integer :: i, k, nacti, nnorb, ntri, ios, nvmono
  integer(c_int64_t) :: n_entries, nval, r0, cnt
  integer(c_int) :: rc
  type(c_ptr) :: ctx
  double precision :: err_est
  character(len=26) :: timeday

  ! ------------------------------------------------------------------ files
  prefix = ' '
  om = 'om'; f25 = 'f25'; f50 = 'f50'; info = 'info'; pqrs = 'pqrs'; ijom = 'ijom'; scr = 'scr'
  pqrs1 = ' '; info1 = ' '; ijom1 = ' '; om1 = ' '; f251 = ' '; f501 = ' '; scr1 = ' '
  read(5, ijfil)
  call fdate(timeday)
  write(6, '(/,80("*"),/,8("   ijkl   "),/,10x,a25,/,80("*"))') timeday
  write(6, '(a)') ' ijkl_gpu: AO->MO transform on MI355X (HIP, FP64 MFMA / split FP16)'
  call open_unit(25, f25, f251, 'replace')
  call open_unit(50, f50, f501, 'replace')
  call open_unit(2, info, info1, 'old')
  call open_unit(8, pqrs, pqrs1, 'old')
  call open_unit(10, ijom, ijom1, 'replace')

  ! ------------------------------------------------------------------ defaults (ijkl.f:37-97)
  nsym = 1; ntrans = -1; istop = 0; yprt = .false.; chstan = .true.; trec = 1.d-5; tdeg = 1.d-6
  inpv = 0; yprtv = .false.; ywvect = .false.; ynocsy = .false.; yrot = .true.; ydp = .false.
  norb = 0; noc = 0; ymono = .false.; yrecsp = .true.; ycv = .false.; yorbf = .true.
  group = ' '; iwant = 0; nveci = 1; nvecf = 0; nvec = 0; ljkf = 0; yhondo = .true.
  yreduc = .true.; nsyma = 0; nsymb = 0; yprth = .false.; yatom = .false.; nrec = 0; mblk = 0
  mijkl = 0; ygel = .false.; itsyv = 0; ivecsy = 0; nocsy = 0; yf11 = .false.; yauto = .true.
  naxis = 2; yuhf = .false.
  do i = 1, nsymz
    isymp(i) = i
  end do
  precision = 0; nslice = 0
  read(5, ijklin, iostat=ios)
  if (ios /= 0) call fail(' erreur sur namelist ijklin', 9)

  ! ------------------------------------------------------------------ info header
  read(2) ioda, nt, nshell, num, nveca, nvecb, bname, ngauss, ne, ich, mul, na, nb, nat, &
          grouph, naxh, title
  if (.not. ydp) call fail(' only YDP=T (double precision integrals) is supported', 9)
  if (yuhf) call fail(' UHF orbitals (two passes) are not supported', 9)
  allocate(ibas(5*nshell), gauss(6*ngauss), atom_c(3*nat), atom_a(nat), atom_ns(nat), &
           atom_ks(nat), atom_iat(nat), atom_zan(nat), newsh(nshell*nt), ptr(9*nt), &
           dtr(36*nt), ftr(100*nt), iflab(3*num), iso(12*nshell))
  rewind 2
  read(2) (ioda(1), i = 1, 19), (k, i = 20, 56), ibas, gauss, &
          (atom_c(3*k-2), atom_c(3*k-1), atom_c(3*k), atom_a(k), atom_ns(k), atom_ks(k), &
           atom_iat(k), atom_zan(k), k = 1, nat), newsh, ptr, dtr, ftr, iflab, iso, invt, &
          enuc, ante
  if (norb == 0) norb = num
  if (norb > doa) call fail(' le nombre d''orbitales depasse la limite du programme', 9)
  if (noc == 0) noc = (ne + 1) / 2
  if (nvecf == 0) nvecf = nveca
  nnorb = norb * norb
  ntri = norb * (norb + 1) / 2
  if (.not. (nt == 1 .and. (grouph(1:2) == 'C1' .or. grouph(1:2) == 'c1'))) then
    call run_symmetric()
    stop
  end if

  if (ntrans == -1) ntrans = 0
  ! ------------------------------------------------------------------ MO coefficients
  ! rotat2 for C1 (rotat2.f:33-60, 71, 386-420): read record 7, compact the active
  ! (not frozen) columns to the front; the symmetry classification and the
  ! normalisation of rotat2.f:72-383 are skipped for a single irreducible representation.
  allocate(c(norb, norb), h(ntri), s(ntri))
  read(2)
  read(2) h
  do i = 1, 3
    read(2)
  end do
  read(2) c
  nacti = 0
  do i = nveci, nvecf
    if (ygel(i)) cycle
    nacti = nacti + 1
    c(:, nacti) = c(:, i)
  end do
  nvec = nvecf - nveci + 1
  do i = 1, nvec
    itsyv(i) = 1
  end do
  write(6, '(a,i5,/,a,i5,/,a,i5,/,a,f20.8)') ' nombre d''orbitales de base', norb, &
      ' nombre d''orbitales moleculaires', nvec, ' orbitales actives (non gelees)', nacti, &
      ' energie nucleaire', enuc

  ! ------------------------------------------------------------------ ijom (ijkl.f:223-240)
  s = 0.d0        ! the S work array of ijkl.f is never filled for one representation
  write(10) c
  write(10) c
  write(10) s
  if (istop == 1) stop

  ! ------------------------------------------------------------------ f25 header (ijkl.f:249-253)
  write(25) nsym, nvec, noc, ntrans, (itsyv(i), i = 1, nvec), norb, ydp, chstan
  write(25) 1

  ! ------------------------------------------------------------------ pqrs stream (itijkl.f:57-68)
  n_entries = int(nrec, c_int64_t) * lsize
  allocate(xx(max(n_entries, 1_c_int64_t)), ix(max(n_entries, 1_c_int64_t)))
  do i = 1, nrec
    r0 = int(i - 1, c_int64_t) * lsize
    read(8) xx(r0 + 1:r0 + lsize), ix(r0 + 1:r0 + lsize)
  end do

  ! ------------------------------------------------------------------ transform on the GPU
  nval = int(nvec, c_int64_t) * (nvec + 1) / 2
  nval = nval * (nval + 1) / 2
  allocate(mo(nval))
  rc = cipsi_ijkl_create(ctx, 0_c_int)
  if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_create: '//cipsi_ijkl_strerror(rc), 9)
  rc = cipsi_ijkl_transform_labels(ctx, int(norb, c_int), int(nvec, c_int), xx, ix, n_entries, &
                                   c, int(norb, c_int), mo, int(precision, c_int), &
                                   int(nslice, c_int), err_est)
  if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_transform_labels: '//cipsi_ijkl_strerror(rc), 9)
  write(6, '(a,i12,a,es10.3)') ' integrales moleculaires:', nval, '  erreur estimee:', err_est

  ! ------------------------------------------------------------------ f50 (tijkld.f:68-88)
  allocate(buf(lsize))
  buf = 0.d0
  mblk = 0
  r0 = 0
  do while (r0 < nval)
    cnt = min(int(lsize, c_int64_t), nval - r0)
    buf(1:cnt) = mo(r0 + 1:r0 + cnt)     ! a short last record keeps the previous tail
    write(50) buf
    mblk = mblk + 1
    r0 = r0 + cnt
  end do
  ! block directory of the single C1 block (itijkl.f:208-213, 291-296)
  write(25) 1, 1, 1
  write(25) 1, int(1, 2)
  write(25) int(1, 2)
  write(25) enuc, ljkf

  ! ------------------------------------------------------------------ tmono (tmono.f:39-56)
  nvmono = nveca
  allocate(hmo(nvmono * (nvmono + 1) / 2))
  rc = cipsi_ijkl_tmono(ctx, int(norb, c_int), int(nvmono, c_int), h, c, int(norb, c_int), hmo)
  if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_tmono: '//cipsi_ijkl_strerror(rc), 9)
  write(50) hmo
  rc = cipsi_ijkl_destroy(ctx)
  close(50)
  close(25)
  close(10)
  call fdate(timeday)
  write(6, '(/,80("*"),/,8(" fin ijkl "),/,10x,a25,/,80("*"))') timeday

contains

  ! The ijkl step for a point group with nsym > 1 (ijkl.f:186-430 with otodon, rotat2,
  ! vector and itijkl in the library).
  subroutine run_symmetric()
    integer(c_int), target :: its(196), jsym(64), isymp_t(nsymz), nbp(nsymz), nbo(nsymz), &
                              ind(nsymz + 1), isydeg_l(28)
    integer(c_int), allocatable, target :: ktype(:), kloc(:), newsh_t(:), iptr(:), idtr(:), &
                                           iftr(:), ishell(:), ktyp(:), numc(:), ivecsy_t(:), &
                                           itsyv_t(:), ygel_t(:), ipair(:), so(:), &
                                           isydeg_t(:)
    double precision, allocatable :: zsh(:)
    integer(c_int16_t), allocatable :: nad(:), ktypsy(:)
    double precision, allocatable, target :: c_out(:), c_packed(:), stream(:)
    character(kind=c_char) :: gname(9)
    integer(c_int) :: nsym_g, nsymp, tag, npaire, max_pairs, ntran
    logical :: linear, dinfh
    integer(c_int64_t) :: count
    type(cipsi_ijkl_sym_t) :: sym
    integer :: sh, j, g, nf, a, nsyt
    integer, parameter :: nfunc(4) = (/ 1, 3, 5, 7 /)

    if (nveci /= 1) call fail(' nveci must be 1', 9)
    if (yuhf) call fail(' UHF orbitals (two passes) are not supported', 9)
    if (itsyv(1) /= 0) call fail(' MO symmetries given in the namelist are not supported', 9)
    if (.not. yauto) call fail(' only yauto=T (tables from the group name) is supported', 9)
    do j = 1, 8
      gname(j) = grouph(j:j)
      if (gname(j) >= 'a' .and. gname(j) <= 'z') gname(j) = achar(iachar(gname(j)) - 32)
    end do
    gname(9) = c_null_char
    dinfh = grouph(1:5) == 'DINFH' .or. grouph(1:5) == 'dinfh'
    linear = dinfh .or. grouph(1:5) == 'CINFV' .or. grouph(1:5) == 'cinfv'
    ntran = 0
    if (.not. linear) then
      ! ---- group tables (otodon.f:163-317)
      rc = cipsi_ijkl_sym_group(gname, int(naxh, c_int), int(nt, c_int), nsym_g, its, jsym)
      if (rc == CIPSI_ERR_UNSUPPORTED) &
        call fail(' programme stoppe : groupe degenere', 9)
      if (rc /= CIPSI_OK) call fail(' groupe non trouve: '//trim(grouph), 9)
      nsym = nsym_g
    end if
    ! ---- basis tables of rondof.f:40-118 (pshf: five pure d, seven pure f functions)
    allocate(ktype(nshell), kloc(nshell), newsh_t(nshell * nt), iptr(3 * nt), idtr(6 * nt), &
             iftr(10 * nt), ishell(norb), ktyp(norb), numc(norb))
    do sh = 1, nshell
      ktype(sh) = ibas(5 * (sh - 1) + 1)
      kloc(sh) = ibas(5 * (sh - 1) + 3)
      if (ktype(sh) < 1 .or. ktype(sh) > 4) &
        call fail(' only s, p, d and f shells of pshf are supported', 9)
    end do
    newsh_t = newsh
    do g = 1, nt
      do j = 1, 3
        iptr(j + 3 * (g - 1)) = int(ptr((3 * (g - 1) + j - 1) * 3 + j))
      end do
      do j = 1, 6
        idtr(j + 6 * (g - 1)) = 0
        if (j <= 5) idtr(j + 6 * (g - 1)) = nint(dtr((6 * (g - 1) + j - 1) * 6 + j))
      end do
      do j = 1, 10
        iftr(j + 10 * (g - 1)) = 0
        if (j <= 7) iftr(j + 10 * (g - 1)) = nint(ftr((10 * (g - 1) + j - 1) * 10 + j))
      end do
    end do
    a = 0
    do sh = 1, nshell
      nf = nfunc(ktype(sh))
      do j = 0, nf - 1
        if (kloc(sh) + j > norb) call fail(' basis tables and N_AO disagree', 9)
        ishell(kloc(sh) + j) = sh
        ktyp(kloc(sh) + j) = ktype(sh)
        numc(kloc(sh) + j) = j
        a = a + 1
      end do
    end do
    if (a /= norb) call fail(' basis tables and N_AO disagree', 9)
    if (linear) then
      ! ---- otodon.f:36-144: classes by angular character about the axis
      allocate(zsh(nshell), so(14 * norb))
      do sh = 1, nshell
        k = ibas(5 * (sh - 1) + 4)                 ! katom
        if (k < 1 .or. k > nat) call fail(' basis tables and atoms disagree', 9)
        zsh(sh) = atom_c(3 * k)
      end do
      rc = cipsi_ijkl_sym_linear(gname, int(norb, c_int), int(nshell, c_int), ktype, kloc, &
                                 zsh, newsh_t, int(nt, c_int), nsym_g, its, ntran, isydeg_l, so)
      if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_sym_linear: '//cipsi_ijkl_strerror(rc), 9)
      nsym = nsym_g
    end if
    ! ijkl.f:105,141-149: ntrans from otodon unless the namelist sets it; one symmetry
    ! transformation per ntrans (yauto), columns past otodon's are zero
    if (ntrans == -1) ntrans = ntran
    allocate(isydeg_t(nsym * max(ntrans, 1)))
    isydeg_t = 0
    do j = 1, min(ntrans, ntran)
      isydeg_t((j - 1) * nsym + 1:j * nsym) = isydeg_l((j - 1) * nsym + 1:j * nsym)
    end do
    ! ---- records 2, 3, 7
    allocate(c(norb, norb), h(ntri), s(ntri))
    read(2) s
    read(2) h
    do j = 1, 3
      read(2)
    end do
    read(2) c
    if (yatom) then                      ! rotat2.f:79-85
      s = 0.d0
      do j = 1, norb
        s(j * (j + 1) / 2) = 1.d0
      end do
    end if
    ! ---- otodon.f:320-353, rotat2.f:50-423, vector.f:17-46
    allocate(c_out(nnorb), c_packed(nnorb), ivecsy_t(nsym * norb), itsyv_t(nvecf), ygel_t(nvecf))
    do j = 1, nvecf
      ygel_t(j) = 0
      if (ygel(j)) ygel_t(j) = 1
    end do
    if (linear) then
      rc = cipsi_ijkl_sym_prepare_so(int(norb, c_int), nsym, so, s, h, c, int(nvecf, c_int), &
                                     ygel_t, int(noc, c_int), trec, tdeg, c_out, itsyv_t, &
                                     isymp_t, nsymp, ivecsy_t, nbp, nbo, ind, c_packed, tag)
    else
      rc = cipsi_ijkl_sym_prepare(int(norb, c_int), nsym, its, jsym, int(nt, c_int), &
                                  int(nshell, c_int), ktype, kloc, newsh_t, iptr, idtr, iftr, &
                                  s, h, c, int(nvecf, c_int), ygel_t, int(noc, c_int), trec, &
                                  tdeg, c_out, itsyv_t, isymp_t, nsymp, ivecsy_t, nbp, nbo, ind, &
                                  c_packed, tag)
    end if
    if (rc == CIPSI_ERR_NOT_ORTHONORMAL) then
      write(6, '(a)') ' rotat2: the MOs do not separate by symmetry or are not orthonormal'
      call fail(' fix ?, error tag =', int(tag))
    end if
    if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_sym_prepare: '//cipsi_ijkl_strerror(rc), 9)
    nvec = nvecf - nveci + 1
    do j = 1, nvec
      itsyv(j) = itsyv_t(j)
    end do
    write(6, '(a,i4)') ' nombre de symetries pour les orb atomiques', nsymp
    write(6, '(a,(1x,30i3))') ' symetrie des om', (itsyv(j), j = 1, nvec)
    write(6, '(a,i5,/,a,i5,/,a,i5,/,a,f20.8)') ' nombre d''orbitales de base', norb, &
        ' nombre d''orbitales moleculaires', nvec, ' nombre de representations irreductibles', &
        nsym, ' energie nucleaire', enuc
    ! ---- ijom (ijkl.f:310-327), f25 header and its (ijkl.f:336-340)
    write(10) c_out
    write(10) c_out
    write(10) s
    if (istop == 1) stop
    write(25) nsym, nvec, noc, ntrans, (itsyv(j), j = 1, nvec), norb, ydp, chstan
    write(25) ((its(a + (j - 1) * nsym), a = 1, nsym), j = 1, nsym)
    if (ntrans /= 0 .and. yhondo) then            ! ijkl.f:254-261
      do j = 1, ntrans
        write(25) 1, (isydeg_t(a + (j - 1) * nsym), a = 1, nsym)
      end do
    end if
    ! ---- pqrs stream and the transform (itijkl.f:57-298)
    n_entries = int(nrec, c_int64_t) * lsize
    allocate(xx(max(n_entries, 1_c_int64_t)), ix(max(n_entries, 1_c_int64_t)))
    do j = 1, nrec
      r0 = int(j - 1, c_int64_t) * lsize
      read(8) xx(r0 + 1:r0 + lsize), ix(r0 + 1:r0 + lsize)
    end do
    nsyt = nsym * (nsym + 1) / 2
    nsyt = min(2000, nsyt * (nsyt + 1) / 2)
    max_pairs = int(nsyt, c_int)
    allocate(ipair(2 * max_pairs), nad(nsyt), ktypsy(nsyt))
    sym%n_ao = int(norb, c_int)
    sym%nsym = nsym
    sym%nt = int(nt, c_int)
    sym%nshell = int(nshell, c_int)
    sym%its = c_loc(its)
    sym%isymp = c_loc(isymp_t)
    sym%nbp = c_loc(nbp)
    sym%nbo = c_loc(nbo)
    sym%ind = c_loc(ind)
    sym%ivecsy = c_loc(ivecsy_t)
    sym%c = c_loc(c_packed)
    sym%ishell = c_loc(ishell)
    sym%kloc = c_loc(kloc)
    sym%ktyp = c_loc(ktyp)
    sym%numc = c_loc(numc)
    sym%newsh = c_loc(newsh_t)
    sym%iptr = c_loc(iptr)
    sym%idtr = c_loc(idtr)
    sym%iftr = c_loc(iftr)
    sym%ntrsy = int(ntrans, c_int)
    sym%isydeg = c_loc(isydeg_t)
    rc = cipsi_ijkl_create(ctx, 0_c_int)
    if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_create: '//cipsi_ijkl_strerror(rc), 9)
    rc = cipsi_ijkl_sym_transform_labels(ctx, sym, xx, ix, n_entries, c_null_ptr, 0_c_int64_t, &
                                         count, ipair, max_pairs, npaire, nad, ktypsy)
    if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_sym_transform_labels: '//cipsi_ijkl_strerror(rc), 9)
    allocate(stream(max(count, 1_c_int64_t)))
    rc = cipsi_ijkl_sym_transform_labels(ctx, sym, xx, ix, n_entries, c_loc(stream), count, &
                                         count, ipair, max_pairs, npaire, nad, ktypsy)
    if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_sym_transform_labels: '//cipsi_ijkl_strerror(rc), 9)
    write(6, '(a,i12,a,i6,a)') ' integrales moleculaires:', count, ' en', npaire, ' blocs'
    ! ---- f50 (tijkld.f:68-88) and the directory (itijkl.f:291-296)
    allocate(buf(lsize))
    buf = 0.d0
    r0 = 0
    do while (r0 < count)
      cnt = min(int(lsize, c_int64_t), count - r0)
      buf(1:cnt) = stream(r0 + 1:r0 + cnt)     ! a short last record keeps the previous tail
      write(50) buf
      r0 = r0 + cnt
    end do
    write(6, *) 'npaire,npair', npaire, 2 * npaire
    write(25) npaire, (ipair(j), j = 1, 2 * npaire)
    write(25) nsyt, (nad(j), j = 1, nsyt)
    write(25) (ktypsy(j), j = 1, nsyt)
    write(25) enuc, ljkf
    ! ---- tmono (ijkl.f:409-423, tmono.f:39-56) on the final coefficients
    nvmono = nveca
    allocate(hmo(nvmono * (nvmono + 1) / 2))
    rc = cipsi_ijkl_tmono(ctx, int(norb, c_int), int(nvmono, c_int), h, c_out, int(norb, c_int), hmo)
    if (rc /= CIPSI_OK) call fail(' cipsi_ijkl_tmono: '//cipsi_ijkl_strerror(rc), 9)
    write(50) hmo
    rc = cipsi_ijkl_destroy(ctx)
    close(50)
    close(25)
    close(10)
    call fdate(timeday)
    write(6, '(/,80("*"),/,8(" fin ijkl "),/,10x,a25,/,80("*"))') timeday
  end subroutine

  ! nomf.f/nomfil.f: file = prefix up to its first blank + root up to its first blank,
  ! unless the full name (second argument) is given.
  subroutine open_unit(iunit, root, full, status)
    integer, intent(in) :: iunit
    character(len=*), intent(in) :: root, full, status
    character(len=180) :: fname
    integer :: ip, ir
    ip = index(prefix, ' ') - 1
    if (ip < 0) ip = len(prefix)
    ir = index(root, ' ') - 1
    if (ir < 0) ir = len(root)
    fname = prefix(1:ip)//root(1:ir)
    if (full /= ' ') fname = full
    open(unit=iunit, file=trim(fname), form='unformatted', status=status, iostat=ios)
    if (ios /= 0) call fail(' cannot open '//trim(fname), 9)
  end subroutine

  subroutine fail(msg, tag)
    character(len=*), intent(in) :: msg
    integer, intent(in) :: tag
    write(6, '(a)') msg
    write(6, '(a17,f15.7)') ' !! ERROR in ijkl', trec
    write(6, *) 'ijkl_gpu, error tag ='
    write(6, '(i2)') tag
    stop 1
  end subroutine

end program ijkl_gpu
