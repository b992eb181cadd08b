!>  Drop-in replacement of the cpw2000 subroutine ditsp_c16
!>  (Src/ditsp/ditsp_c16.f90:20-27): identical dummy arguments, kinds and
!>  intents.  The whole iterative diagonalisation (start from the small
!>  matrix, relaxation, Rayleigh-Ritz, convergence test) runs on the GPU in
!>  libcpwgpu.so through the C ABI declared in include/cpwgpu.h; psi, hpsi and
!>  ei cross PCIe once per call instead of once per H|psi> product.
!>
!>  Link this object instead of Src/ditsp/ditsp_c16.o (with hk_psi_c16.o from
!>  the same directory for the other callers of hk_psi_c16) and add -lcpwgpu.
!>
!>  A nonzero status from the library (the argument checks of
!>  ditsp_c16.f90:151-212, device failures) is reported on unit 6 and ends the
!>  run with STOP, as the reference does.

subroutine ditsp_c16(ipr, ifail, icmax, iguess, epspsi, lnewanl,         &
    neig, mtxd, mtxds,                                                   &
    psi, hpsi, ei,                                                       &
    ekpg, isort, vscr, kmscr,                                            &
    ng, kgv,                                                             &
    anlga, xnlkb, nanl,                                                  &
    hamsm, hdiag,                                                        &
    mxddim, mxdsml, mxdbnd, mxdgve, mxdscr, mxdanl)

  use, intrinsic :: iso_c_binding, only : c_int, c_double, c_double_complex, c_char, c_ptr,   &
                                          c_f_pointer, c_associated

  implicit none

  integer, parameter          :: REAL64 = selected_real_kind(12)

! input

  integer, intent(in)                ::  mxddim                           ! leading dimension of the plane-wave arrays
  integer, intent(in)                ::  mxdsml                           ! leading dimension of the small matrix
  integer, intent(in)                ::  mxdbnd                           ! bound of the band count
  integer, intent(in)                ::  mxdgve                           ! bound of the G list
  integer, intent(in)                ::  mxdscr                           ! length of the potential array
  integer, intent(in)                ::  mxdanl                           ! leading bound of the projector count

  integer, intent(in)                ::  ipr                              ! amount of printing
  integer, intent(in)                ::  icmax                            ! iteration limit
  integer, intent(in)                ::  iguess                           ! 1 when psi holds starting vectors
  real(REAL64), intent(in)           ::  epspsi                           ! target |H psi - e psi| of every band

  integer, intent(in)                ::  mtxd                             ! number of plane waves (rows used)
  integer, intent(in)                ::  mtxds                            ! order of the small starting matrix

  real(REAL64), intent(in)           ::  ekpg(mxddim)                     ! |k+G_i|^2 / 2 (Ha) of row i
  integer, intent(in)                ::  isort(mxddim)                    ! index into kgv of the G of row i
  real(REAL64), intent(in)           ::  vscr(mxdscr)                     ! V(r) on the FFT mesh, index (k3*n2+k2)*id+k1
  integer, intent(in)                ::  kmscr(7)  ! (1:3) potential box, (4:6) mesh n1 n2 n3, (7) leading dim

  integer, intent(in)                ::  ng                               ! length of the G list
  integer, intent(in)                ::  kgv(3,mxdgve)                    ! integer components of each G (reciprocal-lattice units)

  complex(REAL64), intent(in)        ::  anlga(mxddim,mxdanl)             ! projector beta_a(k+G_i): row i, column a
  real(REAL64), intent(in)           ::  xnlkb(mxdanl)                    ! KB strength of each projector
  integer, intent(in)                ::  nanl                             ! number of projectors in use

  complex(REAL64), intent(in)        ::  hamsm(mxdsml,mxdsml)             ! leading mtxds x mtxds block of H
  real(REAL64), intent(in)           ::  hdiag(mxddim)                    ! diagonal of H (preconditioner)

! output

  integer, intent(out)               ::  ifail  ! 0 when converged, otherwise a measure of the residual reached
  real(REAL64), intent(out)          ::  ei(mxdbnd)                       ! eigenvalues (Ha), ascending

! input and output

  complex(REAL64), intent(inout)     ::  psi(mxddim,mxdbnd)               ! bands, one per column
  complex(REAL64), intent(inout)     ::  hpsi(mxddim,mxdbnd)              ! H applied to psi, rows 1..mtxd
  integer, intent(inout)             ::  neig                             ! number of bands
  logical, intent(inout)             ::  lnewanl                          ! projector-change flag: read, never written

  interface
    integer(c_int) function cpwgpu_ditsp_c16(ipr, ifail, icmax, iguess, epspsi, lnewanl,       &
        neig, mtxd, mtxds, psi, hpsi, ei, ekpg, isort, vscr, kmscr, ng, kgv,                    &
        anlga, xnlkb, nanl, hamsm, hdiag,                                                       &
        mxddim, mxdsml, mxdbnd, mxdgve, mxdscr, mxdanl) bind(C, name="cpwgpu_ditsp_c16")
      import :: c_int, c_double, c_double_complex
      integer(c_int), intent(in)                :: ipr, icmax, iguess, lnewanl
      integer(c_int), intent(out)               :: ifail
      real(c_double), intent(in)                :: epspsi
      integer(c_int), intent(inout)             :: neig
      integer(c_int), intent(in)                :: mtxd, mtxds, ng, nanl
      integer(c_int), intent(in)                :: mxddim, mxdsml, mxdbnd, mxdgve, mxdscr, mxdanl
      complex(c_double_complex), intent(inout)  :: psi(*), hpsi(*)
      real(c_double), intent(out)               :: ei(*)
      real(c_double), intent(in)                :: ekpg(*), vscr(*), xnlkb(*), hdiag(*)
      integer(c_int), intent(in)                :: isort(*), kmscr(*), kgv(*)
      complex(c_double_complex), intent(in)     :: anlga(*), hamsm(*)
    end function cpwgpu_ditsp_c16

    type(c_ptr) function cpwgpu_last_error() bind(C, name="cpwgpu_last_error")
      import :: c_ptr
    end function cpwgpu_last_error
  end interface

  integer(c_int)         ::  ist, lnew
  character(len=512)     ::  msg

  lnew = 0
  if(lnewanl) lnew = 1

  ist = cpwgpu_ditsp_c16(ipr, ifail, icmax, iguess, epspsi, lnew,       &
      neig, mtxd, mtxds, psi, hpsi, ei, ekpg, isort, vscr, kmscr,        &
      ng, kgv, anlga, xnlkb, nanl, hamsm, hdiag,                         &
      mxddim, mxdsml, mxdbnd, mxdgve, mxdscr, mxdanl)

  if(ist /= 0) then
    call last_error_text(msg)
    write(6,*)
    write(6,'("  Stopped in ditsp_c16 (GPU):  status ",i4,2x,a)') ist, trim(msg)
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

end subroutine ditsp_c16
