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

  integer, intent(in)                ::  mxddim                           ! leading dimension of the plane-wave arrays
  integer, intent(in)                ::  mxdbnd                           ! bound of the band count
  integer, intent(in)                ::  mxdasp                           ! bound of the spinor projector count
  integer, intent(in)                ::  mxdgve                           ! bound of the G list
  integer, intent(in)                ::  mxdscr                           ! length of the potential array
  integer, intent(in)                ::  mxdnsp                           ! bound of the potential components

  integer, intent(in)                ::  mtxd                             ! number of plane waves (rows used)
  integer, intent(in)                ::  neig                             ! number of bands
  real(REAL64), intent(in)           ::  ekpg(mxddim)                     ! |k+G_i|^2 / 2 (Ha) of row i
  integer, intent(in)                ::  isort(mxddim)                    ! index into kgv of the G of row i

  real(REAL64), intent(in)           ::  vscr_sp(mxdscr,mxdnsp)           ! V(r) components: scalar, sigma_z, sigma_x, sigma_y
  integer, intent(in)                ::  kmscr(7)  ! (1:3) potential box, (4:6) mesh n1 n2 n3, (7) leading dim
  integer, intent(in)                ::  nsp                              ! potential components in use: 1, 2 or 4

  integer, intent(in)                ::  ng                               ! length of the G list
  integer, intent(in)                ::  kgv(3,mxdgve)                    ! integer components of each G (reciprocal-lattice units)

  integer, intent(in)                ::  nanlsp                           ! number of spinor projectors in use
  complex(REAL64), intent(in)        ::  anlsp(2*mxddim,mxdasp)           ! projectors on the interleaved up/down rows
  real(REAL64), intent(in)           ::  xnlkbsp(mxdasp)                  ! KB strength of each spinor projector
  complex(REAL64), intent(in)        ::  psi_sp(2*mxddim,mxdbnd)          ! two-component bands, up/down rows interleaved

! input and output (read only here)

  logical, intent(inout)             ::  lnewanl                          ! projector-change flag: read, never written

! output

  complex(REAL64), intent(out)       ::  hpsi_sp(2*mxddim,mxdbnd)         ! H applied to psi_sp, rows 1..2*mtxd

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
