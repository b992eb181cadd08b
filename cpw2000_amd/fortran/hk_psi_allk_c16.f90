!>  H|psi> for every k-point of a self-consistency step in one call, on the
!>  GPU (libcpwgpu.so, C ABI of include/cpwgpu.h).  The arguments are the all-k
!>  arrays of the k-point loop of cpw_scf (Src/cpw/cpw_scf.f90:488-544:
!>  psi_allk, isort_allk, mtxd_allk, nband) with the per-k kinetic energies and
!>  projectors laid out the same way, and the potential shared by all k-points.
!>  For k-point irk it computes what hk_psi_c16 computes with
!>  mtxd = mtxd_allk(irk), neig = nband(irk), psi = psi_allk(:,:,irk),
!>  ekpg = ekpg_allk(:,irk), isort = isort_allk(:,irk),
!>  anlga = anlga_allk(:,:,irk), xnlkb = xnlkb_allk(:,irk), nanl = nanl_allk(irk).
!>
!>  This subroutine has no counterpart in cpw2000: it batches the k-point loop.
!>  A nonzero status from the library is reported on unit 6 and ends the run
!>  with STOP, as hk_psi_c16 does for the same conditions.

subroutine hk_psi_allk_c16(nrk, mtxd_allk, nband, psi_allk, hpsi_allk,   &
    ng, kgv,                                                             &
    ekpg_allk, isort_allk, vscr, kmscr,                                  &
    anlga_allk, xnlkb_allk, nanl_allk,                                   &
    mxddim, mxdbnd, mxdanl, mxdgve, mxdscr, mxdnrk)

  use, intrinsic :: iso_c_binding, only : c_int, c_double, c_double_complex, c_char, c_ptr,   &
                                          c_f_pointer, c_associated

  implicit none

  integer, parameter          :: REAL64 = selected_real_kind(12)

! input

  integer, intent(in)                ::  mxddim                           ! leading dimension of the plane-wave arrays
  integer, intent(in)                ::  mxdbnd                           ! bound of the band count
  integer, intent(in)                ::  mxdanl                           ! leading bound of the projector count
  integer, intent(in)                ::  mxdgve                           ! bound of the G list
  integer, intent(in)                ::  mxdscr                           ! length of the potential array
  integer, intent(in)                ::  mxdnrk                           ! bound of the k-point count

  integer, intent(in)                ::  nrk                              ! number of k-points
  integer, intent(in)                ::  mtxd_allk(mxdnrk)                ! number of plane waves of each k-point
  integer, intent(in)                ::  nband(mxdnrk)                    ! bands of each k-point
  real(REAL64), intent(in)           ::  ekpg_allk(mxddim,mxdnrk)         ! |k+G_i|^2 / 2 (Ha) of row i of each k-point
  integer, intent(in)                ::  isort_allk(mxddim,mxdnrk)        ! index into kgv of the G of row i of each k-point

  real(REAL64), intent(in)           ::  vscr(mxdscr)                     ! V(r) on the FFT mesh, index (k3*n2+k2)*id+k1
  integer, intent(in)                ::  kmscr(7)  ! (1:3) potential box, (4:6) mesh n1 n2 n3, (7) leading dim

  integer, intent(in)                ::  ng                               ! length of the G list
  integer, intent(in)                ::  kgv(3,mxdgve)                    ! integer components of each G (reciprocal-lattice units)

  integer, intent(in)                ::  nanl_allk(mxdnrk)                ! number of projectors of each k-point
  complex(REAL64), intent(in)        ::  anlga_allk(mxddim,mxdanl,mxdnrk) ! projectors of each k-point, as anlga
  real(REAL64), intent(in)           ::  xnlkb_allk(mxdanl,mxdnrk)        ! KB strengths of each k-point
  complex(REAL64), intent(in)        ::  psi_allk(mxddim,mxdbnd,mxdnrk)   ! bands of each k-point

! output

  complex(REAL64), intent(inout)     ::  hpsi_allk(mxddim,mxdbnd,mxdnrk)  ! H psi of each k-point (other entries kept)

  interface
    integer(c_int) function cpwgpu_hk_psi_allk_c16(nrk, mtxd_allk, nband, psi_allk, hpsi_allk, &
        ng, kgv, ekpg_allk, isort_allk, vscr, kmscr, anlga_allk, xnlkb_allk, nanl_allk,          &
        mxddim, mxdbnd, mxdanl, mxdgve, mxdscr, mxdnrk) bind(C, name="cpwgpu_hk_psi_allk_c16")
      import :: c_int, c_double, c_double_complex
      integer(c_int), intent(in)              :: nrk, ng, mxddim, mxdbnd, mxdanl, mxdgve, mxdscr, mxdnrk
      integer(c_int), intent(in)              :: mtxd_allk(*), nband(*), nanl_allk(*)
      complex(c_double_complex), intent(in)   :: psi_allk(*), anlga_allk(*)
      complex(c_double_complex), intent(inout) :: hpsi_allk(*)
      integer(c_int), intent(in)              :: kgv(*), isort_allk(*), kmscr(*)
      real(c_double), intent(in)              :: ekpg_allk(*), vscr(*), xnlkb_allk(*)
    end function cpwgpu_hk_psi_allk_c16

    type(c_ptr) function cpwgpu_last_error() bind(C, name="cpwgpu_last_error")
      import :: c_ptr
    end function cpwgpu_last_error
  end interface

  integer(c_int)         ::  ist
  character(len=512)     ::  msg

  ist = cpwgpu_hk_psi_allk_c16(nrk, mtxd_allk, nband, psi_allk, hpsi_allk,  &
      ng, kgv, ekpg_allk, isort_allk, vscr, kmscr,                          &
      anlga_allk, xnlkb_allk, nanl_allk,                                    &
      mxddim, mxdbnd, mxdanl, mxdgve, mxdscr, mxdnrk)

  if(ist /= 0) then
    call last_error_text(msg)
    write(6,*)
    write(6,'("   STOPPED in hk_psi_allk_c16 (GPU):  status ",i4,2x,a)') ist, trim(msg)
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

end subroutine hk_psi_allk_c16
