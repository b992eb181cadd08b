!>  Drop-in replacement of the cpw2000 subroutine hk_psi_spin_c16
!>  (Src/hk_psi/reference/hk_psi_spin_c16.f90:28-32): identical dummy
!>  arguments, kinds and intents.  The two-component product H|psi> is
!>  computed on the GPU by libcpwgpu.so through the C ABI declared in
!>  include/cpwgpu.h.
!>
!>  Link this object instead of Src/hk_psi/reference/hk_psi_spin_c16.o and add
!>  -lcpwgpu.  A nonzero status from the library is reported on unit 6 and
!>  ends the run with STOP, as the reference does for the same conditions
!>  (hk_psi_spin_c16.f90:106-141, :176-185).

subroutine hk_psi_spin_c16(mtxd, neig, psi_sp, hpsi_sp, lnewanl,         &
    ng, kgv,                                                             &
    ekpg, isort, vscr_sp, kmscr, nsp,                                    &
    anlsp, xnlkbsp, nanlsp,                                              &
    mxddim, mxdbnd, mxdasp, mxdgve, mxdscr, mxdnsp)

  use, intrinsic :: iso_c_binding, only : c_int, c_double, c_double_complex, c_char, c_ptr,   &
                                          c_f_pointer, c_associated

  implicit none

  integer, parameter          :: REAL64 = selected_real_kind(12)

! input

  integer, intent(in)                ::  mxddim                          !<  array dimension of plane-waves
  integer, intent(in)                ::  mxdbnd                          !<  array dimension for number of bands
  integer, intent(in)                ::  mxdasp                          !<  array dimension of number of projectors
  integer, intent(in)                ::  mxdgve                          !<  array dimension for g-space vectors
  integer, intent(in)                ::  mxdscr                          !<  array dimension of vscr_sp
  integer, intent(in)                ::  mxdnsp                          !<  array dimension for number of spin components

  integer, intent(in)                ::  mtxd                            !<  wavefunction dimension (without spin)
  integer, intent(in)                ::  neig                            !<  number of wavefunctions
  real(REAL64), intent(in)           ::  ekpg(mxddim)                    !<  kinetic energy (hartree) of k+g-vector i
  integer, intent(in)                ::  isort(mxddim)                   !<  g-vector of row/column i

  real(REAL64), intent(in)           ::  vscr_sp(mxdscr,mxdnsp)          !<  potential components (1, sigma_z, sigma_x, sigma_y)
  integer, intent(in)                ::  kmscr(7)                        !<  fft mesh limits and sizes
  integer, intent(in)                ::  nsp                             !<  number of potential components (1,2,4)

  integer, intent(in)                ::  ng                              !<  number of g-vectors
  integer, intent(in)                ::  kgv(3,mxdgve)                   !<  g-vectors in reciprocal lattice coordinates

  integer, intent(in)                ::  nanlsp                          !<  number of projectors
  complex(REAL64), intent(in)        ::  anlsp(2*mxddim,mxdasp)          !<  KB projectors on the spinor rows
  real(REAL64), intent(in)           ::  xnlkbsp(mxdasp)                 !<  KB normalization
  complex(REAL64), intent(in)        ::  psi_sp(2*mxddim,mxdbnd)         !<  spinor wavefunctions

! input and output (read only here)

  logical, intent(inout)             ::  lnewanl                         !<  indicates that anlsp has been recalculated

! output

  complex(REAL64), intent(out)       ::  hpsi_sp(2*mxddim,mxdbnd)        !<  H |psi_sp>

  interface
    integer(c_int) function cpwgpu_hk_psi_spin_c16(mtxd, neig, psi_sp, hpsi_sp, lnewanl,      &
        ng, kgv, ekpg, isort, vscr_sp, kmscr, nsp, anlsp, xnlkbsp, nanlsp,                    &
        mxddim, mxdbnd, mxdasp, mxdgve, mxdscr, mxdnsp) bind(C, name="cpwgpu_hk_psi_spin_c16")
      import :: c_int, c_double, c_double_complex
      integer(c_int), intent(in)              :: mtxd, neig, ng, nsp, nanlsp, lnewanl
      integer(c_int), intent(in)              :: mxddim, mxdbnd, mxdasp, mxdgve, mxdscr, mxdnsp
      complex(c_double_complex), intent(in)   :: psi_sp(*), anlsp(*)
      complex(c_double_complex), intent(out)  :: hpsi_sp(*)
      integer(c_int), intent(in)              :: kgv(*), isort(*), kmscr(*)
      real(c_double), intent(in)              :: ekpg(*), vscr_sp(*), xnlkbsp(*)
    end function cpwgpu_hk_psi_spin_c16

    type(c_ptr) function cpwgpu_last_error() bind(C, name="cpwgpu_last_error")
      import :: c_ptr
    end function cpwgpu_last_error
  end interface

  integer(c_int)         ::  ist, lnew
  character(len=512)     ::  msg

  lnew = 0
  if(lnewanl) lnew = 1

  ist = cpwgpu_hk_psi_spin_c16(mtxd, neig, psi_sp, hpsi_sp, lnew,        &
      ng, kgv, ekpg, isort, vscr_sp, kmscr, nsp, anlsp, xnlkbsp, nanlsp, &
      mxddim, mxdbnd, mxdasp, mxdgve, mxdscr, mxdnsp)

  if(ist /= 0) then
    call last_error_text(msg)
    write(6,*)
    write(6,'("   STOPPED in hk_psi_spin_c16 (GPU):  status ",i4,2x,a)') ist, trim(msg)
    write(6,*)

    stop

  endif

  return

contains

  subroutine last_error_text(txt)
    character(len=*), intent(out)     ::  txt
    type(c_ptr)                       ::  p
    character(kind=c_char), pointer   ::  cs(:)
    integer                           ::  k

    txt = ' '
    p = cpwgpu_last_error()
    if(.not. c_associated(p)) return
    call c_f_pointer(p, cs, [len(txt)])
    do k = 1,len(txt)
      if(cs(k) == achar(0)) exit
      txt(k:k) = cs(k)
    enddo

  end subroutine last_error_text

end subroutine hk_psi_spin_c16
