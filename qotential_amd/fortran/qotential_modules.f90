! Fortran modules of the drop-in boundary: the module names, procedure names and
! explicit-shape argument lists that qotential's Fortran drivers use
! (test/stellarator_bench.f90:2-6; src/lap3d_close_mod.f90:115-133, 622-649,
! 884-960; src/patch_refine_mod.f90:51-55), implemented by ISO_C_BINDING calls
! into libqotential_hip.so.  A driver written against the reference's modules
! compiles unchanged against these and runs its close evaluation on the GPU.
!
! Errors: the library reports failure through qol_last_status / qol_last_error;
! these wrappers print the message and stop, as the reference's `error stop`.

module quatapproximation_mod
  use iso_c_binding, only: c_double
  implicit none
  integer, parameter, public :: r64 = c_double
end module quatapproximation_mod

module qotential_hip_c
  use iso_c_binding
  implicit none
  interface
    function qol_set_level_mode(adaptive, square) bind(C, name='qol_set_level_mode') result(rc)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: adaptive, square
      integer(c_int) :: rc
    end function
    function qol_last_status() bind(C, name='qol_last_status') result(s)
      import :: c_int64_t
      integer(c_int64_t) :: s
    end function
    function qol_last_error() bind(C, name='qol_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    subroutine qol_rrq_mex(m, tx, n, sx, snx, sw, rts, rps, order, nquad, orderff, distff, exterior, &
                           isimd, ichart, sxbd_chart, rv_chart, tgl, wgl, Dgl, w_bclag, Legmat, umatr, &
                           vmatr, As, Ad, Omega, Ias, timeinfo) bind(C, name='qol_rrq_mex_')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: m, n, order, nquad, orderff, exterior, isimd, ichart
      real(c_double) :: tx(*), sx(*), snx(*), sw(*), rts(*), rps(*), distff, sxbd_chart(*), rv_chart(*)
      real(c_double) :: tgl(*), wgl(*), Dgl(*), w_bclag(*), Legmat(*), umatr(*), vmatr(*)
      real(c_double) :: As(*), Ad(*), Omega(*), Ias(*), timeinfo(*)
    end subroutine
    subroutine qol_simplex_precomp_mex(nquad, korder, kpols, tgl, wgl, Dgl, w_bclag, Legmat, umatr, &
                                       vmatr) bind(C, name='qol_simplex_precomp_mex_')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: nquad, korder, kpols
      real(c_double) :: tgl(*), wgl(*), Dgl(*), w_bclag(*), Legmat(*), umatr(*), vmatr(*)
    end subroutine
    subroutine qol_getnearquad_lap_rrq_mex(npan, nterms, hdim, nquad, nsrc, sx, snx, sw, rts, rps, ntcx, &
                                           tcx, tcxrow, tcxi, orderff, distff, iside, isimd, nnz, sparsei, &
                                           sparsej, sparsevs, sparsevd, timeinfo) &
                                           bind(C, name='qol_getnearquad_lap_rrq_mex_')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: npan, nterms, hdim, nquad, nsrc, ntcx, tcxrow(*), tcxi(*), orderff, iside, isimd, nnz
      integer(c_int64_t) :: sparsei(*), sparsej(*)
      real(c_double) :: sx(*), snx(*), sw(*), rts(*), rps(*), tcx(*), distff, sparsevs(*), sparsevd(*), timeinfo(*)
    end subroutine
    subroutine qol_lap3ddlp_closepanel_mex(m_tgt, t_x, npat, s_x, order, ref, if_adapt_d, Ac) &
        bind(C, name='qol_lap3ddlp_closepanel_mex_')
      import :: c_int64_t, c_double
      integer(c_int64_t) :: m_tgt, npat, order, ref
      real(c_double) :: t_x(*), s_x(*), if_adapt_d, Ac(*)
    end subroutine
    function qol_geo_plan_charts(mp, np, p, nfp, nmode, mnrz, iterate_aspect, restol, cap, charts) &
        bind(C, name='qol_geo_plan_charts') result(n)
      import :: c_int64_t, c_double
      integer(c_int64_t), value :: mp, np, p, nfp, nmode, iterate_aspect, cap
      real(c_double), value :: restol
      real(c_double) :: mnrz(*), charts(*)
      integer(c_int64_t) :: n
    end function
    function qol_geo_build_charts(mp, np, p, nfp, nmode, mnrz, nchart, charts, sx, snx, sw, rts, rps) &
        bind(C, name='qol_geo_build_charts') result(rc)
      import :: c_int64_t, c_double, c_int
      integer(c_int64_t), value :: mp, np, p, nfp, nmode, nchart
      real(c_double) :: mnrz(*), charts(*), sx(*), snx(*), sw(*), rts(*), rps(*)
      integer(c_int) :: rc
    end function
  end interface
contains
  subroutine qol_check(what)
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (qol_last_status() == 0_c_int64_t) return
    call c_f_pointer(qol_last_error(), msg, [4096])
    k = 0
    do while (k < 4096)
      if (msg(k + 1) == c_null_char) exit
      k = k + 1
    end do
    write(*, '(a, ": ", 4096a)') what, msg(1:k)
    error stop 'qotential HIP library call failed'
  end subroutine qol_check
end module qotential_hip_c

module patch_refine_mod
  implicit none
  ! the reference's level-construction switches (src/patch_refine_mod.f90:51-55):
  ! square levels (the default), triangle levels (PR_SQUARE = .false.) and the
  ! adaptive walk over triangle levels (PR_ADAPTIVE = .true.); every drop-in call
  ! passes the current values to the library (qol_set_level_mode)
  logical, save :: PR_ADAPTIVE = .false.
  logical, save :: PR_SQUARE = .true.
end module patch_refine_mod

module lap3d_close_mod
  use quatapproximation_mod, only: r64
  use qotential_hip_c
  implicit none
  private
  public :: rrq_r64, simplex_precomp_r64, getnearquad_lap_rrq_r64, Lap3dDLP_closepanel_r64
contains

  ! the module variables to the library; both set stops, as the reference does
  ! (src/patch_refine_mod.f90:116-117)
  subroutine pass_level_mode()
    use patch_refine_mod, only: PR_ADAPTIVE, PR_SQUARE
    integer(c_int) :: rc
    if (PR_ADAPTIVE .and. PR_SQUARE) error stop 'PR_SQUARE and PR_ADAPTIVE cannot be combined'
    rc = qol_set_level_mode(merge(1_c_int64_t, 0_c_int64_t, PR_ADAPTIVE), merge(1_c_int64_t, 0_c_int64_t, PR_SQUARE))
    call qol_check('qol_set_level_mode')
  end subroutine pass_level_mode

  ! src/lap3d_close_mod.f90:115-133
  subroutine simplex_precomp_r64(nquad, korder, kpols, tgl, wgl, Dgl, w_bclag, Legmat, umatr, vmatr)
    integer(8), intent(in) :: nquad, korder, kpols
    real(r64), intent(out) :: tgl(nquad), wgl(nquad)
    real(r64), intent(out) :: Dgl(nquad,nquad), w_bclag(nquad)
    real(r64), intent(out) :: Legmat(nquad,nquad)
    real(r64), intent(out) :: umatr(kpols,kpols), vmatr(kpols,kpols)
    integer(c_int64_t) :: nq, ko, kp
    nq = nquad; ko = korder; kp = kpols
    call qol_simplex_precomp_mex(nq, ko, kp, tgl, wgl, Dgl, w_bclag, Legmat, umatr, vmatr)
    call qol_check('simplex_precomp_r64')
  end subroutine simplex_precomp_r64

  ! src/lap3d_close_mod.f90:622-649 (exterior is a logical here, as in the reference)
  subroutine rrq_r64(m, tx, n, sx, snx, sw, rts, rps, order, nquad, orderff, distff, exterior, isimd, &
                     ichart, sxbd_chart, rv_chart, tgl, wgl, Dgl, w_bclag, Legmat, umatr, vmatr, &
                     As, Ad, Omega, IalphaAsvestas, timeinfo)
    integer(8), intent(in)    :: m, n, order, nquad, orderff, isimd
    real(r64),  intent(in)    :: tx(3,m), sx(3,n), snx(3,n), sw(n)
    real(r64),  intent(in)    :: rts(3,n), rps(3,n), distff
    logical,    intent(in)    :: exterior
    real(r64),  intent(out)   :: As(m,n), Ad(m,n)
    real(r64),  intent(out)   :: Omega(4*(order*(order+1)/2),m)
    real(r64),  intent(out)   :: IalphaAsvestas(m)
    real(r64),  intent(inout) :: timeinfo(20)
    integer(8), intent(in)    :: ichart
    real(r64),  intent(in)    :: sxbd_chart(3,3*nquad), rv_chart(3,3)
    real(r64),  intent(in)    :: tgl(nquad), wgl(nquad)
    real(r64),  intent(in)    :: Dgl(nquad,nquad), w_bclag(nquad)
    real(r64),  intent(in)    :: Legmat(nquad,nquad)
    real(r64),  intent(in)    :: umatr(order*(order+1)/2,order*(order+1)/2), &
                                 vmatr(order*(order+1)/2,order*(order+1)/2)
    integer(c_int64_t) :: cm, cn, co, cq, cf, ce, cs, ci
    real(c_double) :: cd
    call pass_level_mode()
    cm = m; cn = n; co = order; cq = nquad; cf = orderff; cs = isimd; ci = ichart; cd = distff
    ce = merge(1_c_int64_t, 0_c_int64_t, exterior)
    call qol_rrq_mex(cm, tx, cn, sx, snx, sw, rts, rps, co, cq, cf, cd, ce, cs, ci, sxbd_chart, rv_chart, &
                     tgl, wgl, Dgl, w_bclag, Legmat, umatr, vmatr, As, Ad, Omega, IalphaAsvestas, timeinfo)
    call qol_check('rrq_r64')
  end subroutine rrq_r64

  ! src/lap3d_close_mod.f90:884-960
  subroutine getnearquad_lap_rrq_r64(npan, nterms, hdim, nquad, nsrc, sx, snx, sw, rts, rps, ntcx, tcx, &
                                     tcxrow, tcxi, orderff, distff, iside, isimd, nnz, sparsei, sparsej, &
                                     sparsevs, sparsevd, timeinfo)
    integer(8), intent(in)  :: npan, nterms, hdim, nquad, nsrc, ntcx, orderff, iside, isimd, nnz
    real(r64),  intent(in)  :: sx(3,nsrc), snx(3,nsrc), sw(nsrc), rts(3,nsrc), rps(3,nsrc)
    real(r64),  intent(in)  :: tcx(3,ntcx), distff
    integer(8), intent(in)  :: tcxrow(ntcx), tcxi(npan+1)
    integer(8), intent(out) :: sparsei(nnz), sparsej(nnz)
    real(r64),  intent(out) :: sparsevs(nnz), sparsevd(nnz)
    real(r64),  intent(inout) :: timeinfo(20)
    integer(c_int64_t) :: a1, a2, a3, a4, a5, a6, a7, a8, a9, a10
    real(c_double) :: cd
    call pass_level_mode()
    a1 = npan; a2 = nterms; a3 = hdim; a4 = nquad; a5 = nsrc; a6 = ntcx; a7 = orderff; a8 = iside
    a9 = isimd; a10 = nnz; cd = distff
    call qol_getnearquad_lap_rrq_mex(a1, a2, a3, a4, a5, sx, snx, sw, rts, rps, a6, tcx, tcxrow, tcxi, a7, cd, &
                                     a8, a9, a10, sparsei, sparsej, sparsevs, sparsevd, timeinfo)
    call qol_check('getnearquad_lap_rrq_r64')
  end subroutine getnearquad_lap_rrq_r64

  ! src/lap3d_close_mod.f90:135-141
  subroutine Lap3dDLP_closepanel_r64(m_tgt, t_x, npat, s_x, order, ref, if_adapt, Ac)
    integer(8), intent(in)  :: m_tgt, npat, order, ref
    real(r64),  intent(in)  :: t_x(3, m_tgt)
    real(r64),  intent(in)  :: s_x(3, npat)
    logical,    intent(in)  :: if_adapt
    real(r64),  intent(out) :: Ac(m_tgt, npat)
    integer(c_int64_t) :: a1, a2, a3, a4
    real(c_double) :: fa
    a1 = m_tgt; a2 = npat; a3 = order; a4 = ref
    fa = merge(1.0_c_double, 0.0_c_double, if_adapt)
    call qol_lap3ddlp_closepanel_mex(a1, t_x, a2, s_x, a3, a4, fa, Ac)
    call qol_check('Lap3dDLP_closepanel_r64')
  end subroutine Lap3dDLP_closepanel_r64

end module lap3d_close_mod

module stellarator_mesh_mod
  use quatapproximation_mod, only: r64
  use qotential_hip_c
  implicit none
  private
  public :: stellarator_mesh_init_r64, create_stellarator_tri_mesh_r64
contains

  ! VMEC mode rows (m, n, Rc, Zs) in the library's layout mnrz(4, nmode)
  subroutine pack_modes(nmode, mn, rc, zs, mnrz)
    integer(8), intent(in) :: nmode, mn(2,*)
    real(r64),  intent(in) :: rc(*), zs(*)
    real(r64),  intent(out) :: mnrz(4, max(nmode, 1_8))
    integer(8) :: k
    mnrz = 0.0_r64
    do k = 1, nmode
      mnrz(1,k) = real(mn(1,k), r64); mnrz(2,k) = real(mn(2,k), r64)
      mnrz(3,k) = rc(k); mnrz(4,k) = zs(k)
    end do
  end subroutine pack_modes

  ! the chart list of a stellarator surface: nmode = 0 the built-in surface
  ! (one aspect pass), else the VMEC modes (aspect splits to convergence, and
  ! curvature refinement with 2:1 balance for restol > 0); ntri = 2 nchart
  subroutine stellarator_mesh_init_r64(mp, np, nterms, nfp, nmode, mn, rc, zs, restol, cap, charts, &
                                       nchart, ntri, ier)
    integer(8), intent(in)  :: mp, np, nterms, nfp, nmode, cap
    integer(8), intent(in)  :: mn(2,*)
    real(r64),  intent(in)  :: rc(*), zs(*), restol
    real(r64),  intent(out) :: charts(6,cap)
    integer(8), intent(out) :: nchart, ntri, ier
    real(r64) :: mnrz(4, max(nmode, 1_8))
    integer(c_int64_t) :: n, it
    call pack_modes(nmode, mn, rc, zs, mnrz)
    it = merge(0_c_int64_t, 1_c_int64_t, nmode == 0)
    n = qol_geo_plan_charts(mp, np, nterms, nfp, nmode, mnrz, it, restol, cap, charts)
    ier = 0
    nchart = 0
    ntri = 0
    if (n < 0) then
      ier = 1
      return
    end if
    nchart = n
    ntri = 2 * n
  end subroutine stellarator_mesh_init_r64

  subroutine create_stellarator_tri_mesh_r64(mp, np, nterms, nfp, nmode, mn, rc, zs, nchart, charts, ntri, &
                                             sx, snx, sw, rts, rps, ier)
    integer(8), intent(in)  :: mp, np, nterms, nfp, nmode, nchart, ntri
    integer(8), intent(in)  :: mn(2,*)
    real(r64),  intent(in)  :: rc(*), zs(*), charts(6,nchart)
    real(r64),  intent(out) :: sx(3,*), snx(3,*), sw(*), rts(3,*), rps(3,*)
    integer(8), intent(out) :: ier
    real(r64) :: mnrz(4, max(nmode, 1_8))
    call pack_modes(nmode, mn, rc, zs, mnrz)
    ier = 0
    if (ntri /= 2 * nchart) then
      ier = 1
      return
    end if
    if (qol_geo_build_charts(mp, np, nterms, nfp, nmode, mnrz, nchart, charts, sx, snx, sw, rts, rps) /= 0) ier = 1
  end subroutine create_stellarator_tri_mesh_r64

end module stellarator_mesh_mod
