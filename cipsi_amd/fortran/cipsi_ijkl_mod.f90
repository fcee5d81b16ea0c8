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
