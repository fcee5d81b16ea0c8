! ISO_C_BINDING interface to the C ABI of libcipsi_ijkl.so (include/cipsi_ijkl.h).
!
! Fortran callers pass their own arrays; the library never retains a pointer and
! every function returns an integer status (CIPSI_OK = 0, negative on error).
module cipsi_ijkl_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CIPSI_OK = 0
  integer(c_int), parameter :: CIPSI_ERR_ARG = -1
  integer(c_int), parameter :: CIPSI_ERR_LABEL = -2
  integer(c_int), parameter :: CIPSI_ERR_NOMEM = -3
  integer(c_int), parameter :: CIPSI_ERR_HIP = -4
  integer(c_int), parameter :: CIPSI_ERR_NONFINITE = -5
  integer(c_int), parameter :: CIPSI_ERR_NOT_ORTHONORMAL = -6
  integer(c_int), parameter :: CIPSI_ERR_UNSUPPORTED = -7
  integer(c_int), parameter :: CIPSI_ERR_NODEVICE = -8
  integer(c_int), parameter :: CIPSI_PREC_FP64 = 0
  integer(c_int), parameter :: CIPSI_PREC_SPLIT_FP16 = 1

  ! cipsi_ijkl_sym: the COMMON state itijkl reads for a molecule with symmetry (pointers
  ! to arrays of 1-based values; include/cipsi_ijkl.h documents each member)
  type, bind(C) :: cipsi_ijkl_sym_t
    integer(c_int) :: n_ao, nsym, nt, nshell
    type(c_ptr) :: its, isymp, nbp, nbo, ind, ivecsy, c
    type(c_ptr) :: ishell, kloc, ktyp, numc, newsh, iptr, idtr, iftr
    integer(c_int) :: ntrsy = 0
    type(c_ptr) :: isydeg = c_null_ptr
  end type

  interface
    integer(c_int) function cipsi_ijkl_create(ctx, device) bind(C, name="cipsi_ijkl_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function

    integer(c_int) function cipsi_ijkl_destroy(ctx) bind(C, name="cipsi_ijkl_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function cipsi_ijkl_strerror_c(status) bind(C, name="cipsi_ijkl_strerror")
      import :: c_int, c_ptr
      integer(c_int), value :: status
    end function

    integer(c_int) function cipsi_ijkl_transform_labels(ctx, n_ao, n_mo, xx, labels, n_entries, &
        c, ldc, mo_packed, precision, n_slices, err_est) bind(C, name="cipsi_ijkl_transform_labels")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_ao, n_mo, ldc, precision, n_slices
      real(c_double), intent(in) :: xx(*)
      integer(c_int32_t), intent(in) :: labels(*)
      integer(c_int64_t), value :: n_entries
      real(c_double), intent(in) :: c(*)
      real(c_double), intent(out) :: mo_packed(*)
      real(c_double), intent(out) :: err_est
    end function

    integer(c_int) function cipsi_ijkl_transform_packed(ctx, n_ao, n_mo, ao_packed, c, ldc, &
        mo_packed, precision, n_slices, err_est) bind(C, name="cipsi_ijkl_transform_packed")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_ao, n_mo, ldc, precision, n_slices
      real(c_double), intent(in) :: ao_packed(*), c(*)
      real(c_double), intent(out) :: mo_packed(*)
      real(c_double), intent(out) :: err_est
    end function

    integer(c_int) function cipsi_ijkl_tmono(ctx, n_ao, n_mo, h_packed, c, ldc, h_mo) &
        bind(C, name="cipsi_ijkl_tmono")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_ao, n_mo, ldc
      real(c_double), intent(in) :: h_packed(*), c(*)
      real(c_double), intent(out) :: h_mo(*)
    end function

    ! E(C) = sum h^MO gamma + 1/2 sum Gamma (pq|rs)_MO and dE/dC (n_ao x n_mo).
    ! gamma is n_mo x n_mo; Gamma_packed is packed like the MO stream.
    integer(c_int) function cipsi_ijkl_energy_grad(ctx, n_ao, n_mo, ao_packed, h_packed, c, &
        ldc, gamma, Gamma_packed, energy, grad) bind(C, name="cipsi_ijkl_energy_grad")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_ao, n_mo, ldc
      real(c_double), intent(in) :: ao_packed(*), h_packed(*), c(*), gamma(*), Gamma_packed(*)
      real(c_double), intent(out) :: energy, grad(*)
    end function

    ! A whole scan in one call: n_geom device pointers per argument (FP64 only; the
    ! context keeps the plan of its last size list and replays its graph for the same list).
    integer(c_int) function cipsi_ijkl_transform_batch_dev(ctx, n_geom, n_ao, n_mo, d_ao, d_c, &
        ldc, d_mo, precision, n_slices) bind(C, name="cipsi_ijkl_transform_batch_dev")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_geom, precision, n_slices
      integer(c_int), intent(in) :: n_ao(*), n_mo(*), ldc(*)
      type(c_ptr), intent(in) :: d_ao(*), d_c(*), d_mo(*)
    end function

    ! Closed-shell energy (no nuclear term), Fock (|F| < 1e-7 -> 0), J and K of
    ! fock.f:80-127 for noc doubly occupied MOs; all packed i >= j like h_mo_packed.
    integer(c_int) function cipsi_ijkl_fock(ctx, n_mo, noc, mo_packed, h_mo_packed, energy, &
        fock, jint, kint) bind(C, name="cipsi_ijkl_fock")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_mo, noc
      real(c_double), intent(in) :: mo_packed(*), h_mo_packed(*)
      real(c_double), intent(out) :: energy, fock(*), jint(*), kint(*)
    end function

    ! cip's hdig / hntd (cip/hdig.f, cip/hntd.f) for batches of determinants held in
    ! cip's /spo/ arrays (ne, nd, trou, part, iorb, ispin; 1-based numbers), all device
    ! buffers; out(d) = E(D_d) - E(reference), out(q) = <D_a|H|D_b> for pairs(2q-1:2q).
    integer(c_int) function cipsi_ijkl_hdig_dev(ctx, n_mo, n_det, n_exc, ne, nd, trou, part, &
        iorb, ispin, fock, jint, kint, out) bind(C, name="cipsi_ijkl_hdig_dev")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, ne, nd, trou, part, iorb, ispin, fock, jint, kint, out
      integer(c_int), value :: n_mo, n_det
      integer(c_int64_t), value :: n_exc
    end function
    integer(c_int) function cipsi_ijkl_hntd_dev(ctx, n_mo, n_det, n_exc, ne, nd, trou, part, &
        iorb, ispin, pairs, n_pairs, mo_packed, fock, out) bind(C, name="cipsi_ijkl_hntd_dev")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, ne, nd, trou, part, iorb, ispin, pairs, mo_packed, fock, out
      integer(c_int), value :: n_mo, n_det
      integer(c_int64_t), value :: n_exc, n_pairs
    end function
    ! Point-group symmetry: group tables (otodon.f:163-317), the MO preparation of
    ! otodon / rotat2 / vector (error tags of rotat2 in tag), and the symmetric itijkl.
    integer(c_int) function cipsi_ijkl_sym_group(group, naxis, nt, nsym, its, jsym) &
        bind(C, name="cipsi_ijkl_sym_group")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: group(*)
      integer(c_int), value :: naxis, nt
      integer(c_int), intent(out) :: nsym, its(*), jsym(*)
    end function
    integer(c_int) function cipsi_ijkl_sym_prepare(n_ao, nsym, its, jsym, nt, nshell, ktype, &
        kloc, newsh, iptr, idtr, iftr, s_packed, h_packed, c_square, nvecf, ygel, noc, trec, &
        tdeg, c_out, itsyv, isymp, nsymp, ivecsy, nbp, nbo, ind, c_packed, tag) &
        bind(C, name="cipsi_ijkl_sym_prepare")
      import :: c_int, c_double
      integer(c_int), value :: n_ao, nsym, nt, nshell, nvecf, noc
      integer(c_int), intent(in) :: its(*), jsym(*), ktype(*), kloc(*), newsh(*), iptr(*), &
                                    idtr(*), iftr(*), ygel(*)
      real(c_double), intent(in) :: s_packed(*), h_packed(*), c_square(*)
      real(c_double), value :: trec, tdeg
      real(c_double), intent(out) :: c_out(*), c_packed(*)
      integer(c_int), intent(out) :: itsyv(*), isymp(*), nsymp, ivecsy(*), nbp(*), nbo(*), &
                                     ind(*), tag
    end function
    ! The linear groups CINFV / DINFH: otodon.f:36-144 (tables and symmetry orbitals from
    ! the basis), then rotat2 / vector on otodon's table.
    integer(c_int) function cipsi_ijkl_sym_linear(group, n_ao, nshell, ktype, kloc, zsh, &
        newsh, nt, nsym, its, ntrans, isydeg, so) bind(C, name="cipsi_ijkl_sym_linear")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: group(*)
      integer(c_int), value :: n_ao, nshell, nt
      integer(c_int), intent(in) :: ktype(*), kloc(*), newsh(*)
      real(c_double), intent(in) :: zsh(*)
      integer(c_int), intent(out) :: nsym, its(*), ntrans, isydeg(*), so(*)
    end function
    integer(c_int) function cipsi_ijkl_sym_prepare_so(n_ao, nsym, so, s_packed, h_packed, &
        c_square, nvecf, ygel, noc, trec, tdeg, c_out, itsyv, isymp, nsymp, ivecsy, nbp, nbo, &
        ind, c_packed, tag) bind(C, name="cipsi_ijkl_sym_prepare_so")
      import :: c_int, c_double
      integer(c_int), value :: n_ao, nsym, nvecf, noc
      integer(c_int), intent(in) :: so(*), ygel(*)
      real(c_double), intent(in) :: s_packed(*), h_packed(*), c_square(*)
      real(c_double), value :: trec, tdeg
      real(c_double), intent(out) :: c_out(*), c_packed(*)
      integer(c_int), intent(out) :: itsyv(*), isymp(*), nsymp, ivecsy(*), nbp(*), nbo(*), &
                                     ind(*), tag
    end function
    integer(c_int) function cipsi_ijkl_sym_transform_labels(ctx, sym, xx, labels, n_entries, &
        mo_stream, mo_capacity, mo_count, ipair, max_pairs, npaire, nad, ktypsy) &
        bind(C, name="cipsi_ijkl_sym_transform_labels")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t, c_int16_t, cipsi_ijkl_sym_t
      type(c_ptr), value :: ctx
      type(cipsi_ijkl_sym_t), intent(in) :: sym
      real(c_double), intent(in) :: xx(*)
      integer(c_int32_t), intent(in) :: labels(*)
      integer(c_int64_t), value :: n_entries, mo_capacity
      type(c_ptr), value :: mo_stream            ! c_null_ptr: sizes and directory only
      integer(c_int64_t), intent(out) :: mo_count
      integer(c_int), intent(out) :: ipair(*), npaire
      integer(c_int), value :: max_pairs
      integer(c_int16_t), intent(out) :: nad(*), ktypsy(*)
    end function
  end interface

contains

  ! Message text of a status code as a Fortran string.
  function cipsi_ijkl_strerror(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    cp = cipsi_ijkl_strerror_c(status)
    call c_f_pointer(cp, p, [256])
    n = 0
    do while (n < 256)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module cipsi_ijkl_mod
