!>  Drop-in replacement of the cpw2000 subroutine hk_psi_c16
!>  (Src/hk_psi/reference/hk_psi_c16.f90:28-78): identical dummy arguments,
!>  kinds and intents.  The product H|psi> is computed on the GPU by
!>  libcpwgpu.so, reached through the C ABI declared in include/cpwgpu.h.
!>
!>  Link this object instead of Src/hk_psi/reference/hk_psi_c16.o (the link-time
!>  selection of Src/Makefile:76-82) and add -lcpwgpu.
!>
!>  A nonzero status from the library is reported on unit 6 and ends the run
!>  with STOP, as the reference does for the same conditions
!>  (hk_psi_c16.f90:108, :112-119, :162-172).

subroutine hk_psi_c16(mtxd, neig, psi, hpsi, lnewanl,                    &
    ng, kgv,                                                             &
    ekpg, isort, vscr, kmscr,                                            &
    anlga, xnlkb, nanl,                                                  &
    mxddim, mxdbnd, mxdanl, mxdgve, mxdscr)

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

  integer, intent(in)                ::  mtxd                             ! number of plane waves (rows used)
  integer, intent(in)                ::  neig                             ! number of bands
  real(REAL64), intent(in)           ::  ekpg(mxddim)                     ! |k+G_i|^2 / 2 (Ha) of row i
  integer, intent(in)                ::  isort(mxddim)                    ! index into kgv of the G of row i

  real(REAL64), intent(in)           ::  vscr(mxdscr)                     ! V(r) on the FFT mesh, index (k3*n2+k2)*id+k1
  integer, intent(in)                ::  kmscr(7)  ! (1:3) potential box, (4:6) mesh n1 n2 n3, (7) leading dim

  integer, intent(in)                ::  ng                               ! length of the G list
  integer, intent(in)                ::  kgv(3,mxdgve)                    ! integer components of each G (reciprocal-lattice units)

  integer, intent(in)                ::  nanl                             ! number of projectors in use
  complex(REAL64), intent(in)        ::  anlga(mxddim,mxdanl)             ! projector beta_a(k+G_i): row i, column a
  real(REAL64), intent(in)           ::  xnlkb(mxdanl)                    ! KB strength of each projector
  complex(REAL64), intent(in)        ::  psi(mxddim,mxdbnd)               ! bands, one per column

! input and output (read only here: some callers pass a literal)

  logical, intent(inout)             ::  lnewanl                          ! projector-change flag: read, never written

! output

  complex(REAL64), intent(out)       ::  hpsi(mxddim,mxdbnd)              ! H applied to psi, rows 1..mtxd

  interface
    integer(c_int) function cpwgpu_hk_psi_c16(mtxd, neig, psi, hpsi, lnewanl,   &
        ng, kgv, ekpg, isort, vscr, kmscr, anlga, xnlkb, nanl,                    &
        mxddim, mxdbnd, mxdanl, mxdgve, mxdscr) bind(C, name="cpwgpu_hk_psi_c16")
      import :: c_int, c_double, c_double_complex
      integer(c_int), intent(in)              :: mtxd, neig, ng, nanl
      integer(c_int), intent(in)              :: mxddim, mxdbnd, mxdanl, mxdgve, mxdscr
      complex(c_double_complex), intent(in)   :: psi(*)
      complex(c_double_complex), intent(out)  :: hpsi(*)
      integer(c_int), intent(in)              :: lnewanl
      integer(c_int), intent(in)              :: kgv(*), isort(*), kmscr(*)
      real(c_double), intent(in)              :: ekpg(*), vscr(*), xnlkb(*)
      complex(c_double_complex), intent(in)   :: anlga(*)
    end function cpwgpu_hk_psi_c16

    type(c_ptr) function cpwgpu_last_error() bind(C, name="cpwgpu_last_error")
      import :: c_ptr
    end function cpwgpu_last_error
  end interface

  integer(c_int)         ::  ist, lnew
  character(len=512)     ::  msg

  lnew = 0
  if(lnewanl) lnew = 1

  ist = cpwgpu_hk_psi_c16(mtxd, neig, psi, hpsi, lnew,                   &
      ng, kgv, ekpg, isort, vscr, kmscr, anlga, xnlkb, nanl,             &
      mxddim, mxdbnd, mxdanl, mxdgve, mxdscr)

  if(ist /= 0) then
    call last_error_text(msg)
    write(6,*)
    write(6,'("   STOPPED in hk_psi_c16 (GPU):  status ",i4,2x,a)') ist, trim(msg)
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

end subroutine hk_psi_c16
