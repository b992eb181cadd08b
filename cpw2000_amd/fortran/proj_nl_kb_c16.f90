!>  Kleinman-Bylander projectors of one k-point computed on the GPU
!>  (libcpwgpu.so, C ABI of include/cpwgpu.h).  Same dummy arguments, kinds
!>  and intents as the reference's proj_nl_kb_c16
!>  (Src/omp/proj_nl_kb_c16.f90:23-73); a build links this file in its place.
!>  A nonzero status from the library is reported on unit 6 and ends the run
!>  with STOP, as the reference does for its inconsistent-nanl, mxdanl and
!>  cell checks.

subroutine proj_nl_kb_c16(rkpt, mtxd, isort, nanl,                       &
    ng, kgv,                                                             &
    nqnl, delqnl, vkb, nkb,                                              &
    ntype, natom, rat, adot,                                             &
    anlga, xnlkb,                                                        &
    mxdtyp, mxdatm, mxdlqp, mxddim, mxdanl, mxdgve)

  use, intrinsic :: iso_c_binding, only : c_int, c_double, c_double_complex, c_char, c_ptr,   &
                                          c_f_pointer, c_associated

  implicit none

  integer, parameter          :: REAL64 = selected_real_kind(12)
  integer, parameter          :: LMAX = 3

! input

  integer, intent(in)                ::  mxdtyp                           ! bound of the species count
  integer, intent(in)                ::  mxdatm                           ! bound of atoms per species
  integer, intent(in)                ::  mxddim                           ! leading dimension of the plane-wave arrays
  integer, intent(in)                ::  mxdlqp                           ! last index of the vkb tables
  integer, intent(in)                ::  mxdanl                           ! leading bound of the projector count
  integer, intent(in)                ::  mxdgve                           ! bound of the G list

  real(REAL64), intent(in)           ::  rkpt(3)                          ! k-point, reciprocal-lattice coordinates
  integer, intent(in)                ::  mtxd                             ! number of plane waves (rows used)
  integer, intent(in)                ::  isort(mxddim)                    ! index into kgv of the G of row i

  integer, intent(in)                ::  nanl                             ! number of projectors in use
  integer, intent(in)                ::  ng                               ! length of the G list
  integer, intent(in)                ::  kgv(3,mxdgve)                    ! integer components of each G (reciprocal-lattice units)
  integer, intent(in)                ::  ntype                            ! number of species
  integer, intent(in)                ::  natom(mxdtyp)                    ! atoms of each species
  real(REAL64), intent(in)           ::  rat(3,mxdatm,mxdtyp)             ! atom positions, fractional (lattice) coordinates
  real(REAL64), intent(in)           ::  adot(3,3)                        ! real-space metric a_i . a_j (bohr^2)
  integer, intent(in)                ::  nqnl(mxdtyp)                     ! points of the vkb tables of each species
  real(REAL64), intent(in)           ::  delqnl(mxdtyp)                   ! q step of the vkb tables of each species
  real(REAL64), intent(in)        ::  vkb(-2:mxdlqp,0:LMAX,-1:1,mxdtyp)   ! radial KB tables over q (divided by q^l)
  integer, intent(in)                ::  nkb(0:LMAX,-1:1,mxdtyp)          ! sign (+1/-1/0) of each (l, species) KB term

! output

  complex(REAL64), intent(out)       ::  anlga(mxddim,mxdanl)             ! projector beta_a(k+G_i): row i, column a
  real(REAL64), intent(out)          ::  xnlkb(mxdanl)                    ! KB strength of each projector

  interface
    integer(c_int) function cpwgpu_proj_nl_kb_c16(rkpt, mtxd, isort, nanl, ng, kgv, nqnl, delqnl, vkb, nkb, &
        ntype, natom, rat, adot, anlga, xnlkb, mxdtyp, mxdatm, mxdlqp, mxddim, mxdanl, mxdgve)            &
        bind(C, name="cpwgpu_proj_nl_kb_c16")
      import :: c_int, c_double, c_double_complex
      real(c_double), intent(in)     :: rkpt(*), delqnl(*), vkb(*), rat(*), adot(*)
      integer(c_int), intent(in)     :: mtxd, nanl, ng, ntype, mxdtyp, mxdatm, mxdlqp, mxddim, mxdanl, mxdgve
      integer(c_int), intent(in)     :: isort(*), kgv(*), nqnl(*), nkb(*), natom(*)
      complex(c_double_complex), intent(inout) :: anlga(*)
      real(c_double), intent(inout)  :: xnlkb(*)
    end function cpwgpu_proj_nl_kb_c16

    type(c_ptr) function cpwgpu_last_error() bind(C, name="cpwgpu_last_error")
      import :: c_ptr
    end function cpwgpu_last_error
  end interface

  integer(c_int)         ::  ist
  character(len=512)     ::  msg

  ist = cpwgpu_proj_nl_kb_c16(rkpt, mtxd, isort, nanl, ng, kgv, nqnl, delqnl, vkb, nkb,   &
      ntype, natom, rat, adot, anlga, xnlkb, mxdtyp, mxdatm, mxdlqp, mxddim, mxdanl, mxdgve)

  if(ist /= 0) then
    call last_error_text(msg)
    write(6,*)
    write(6,'("   STOPPED in proj_nl_kb_c16 (GPU):  status ",i4,2x,a)') ist, trim(msg)
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

end subroutine proj_nl_kb_c16
