! kmm_dropin.f90 -- Fortran modules named as the reference's (derivs,
! solvers, helpers) whose procedures keep the reference's names and dummy
! argument lists and run on the GPU through the C ABI of kmm.h.
!
! A reference host keeps its calls unchanged after `call kmm_attach(ctx, ny,
! nz)` with a context made by kmm_create (module kmm_c).  Arrays are the
! reference's: spectral fields complex(8) (nyp,nz,nxh), profiles real(8)
! (nyp), one kx slab complex(8) (nyp,nz) for solve/penta.  The reference's
! work arrays (a, wrkc, wrk1) are accepted and not used.  A failing call
! prints the library's message and stops, as the reference stops.
!
! Reference interfaces (FFTW_DNS/code/src/dns):
!   derivs:  cderiv derivatives.f90:258, gradscl :90, derivscji :148
!   solvers: solve solvers.f90:174, penta :680, phirhs :555, nonlin :1497,
!            veloc :1576, vort :1663, uzero :1065, vcw3dp :1726 (the NS,
!            SCALAR and SCALAR+POLYMER argument lists under one generic name;
!            the step index the reference reads from its COMMON block is
!            kmm_it, which the host sets before the call)
!   helpers: norm helpers.f90:91, polyforce :498, subforce :564, polyNL :610
module kmm_dropin_state
  use iso_c_binding
  use kmm_c
  implicit none
  type(c_ptr), save :: kmm_ctx_handle = c_null_ptr
  integer, save :: kmm_it = 0
  ! leading extents (nyp, nz) of the (nyp,nz,nxh) fields that vcw3dp's
  ! rank-3 dummies take (a generic name resolves by rank)
  integer, save :: kmm_nyp = 1, kmm_nz = 1
contains
  subroutine kmm_attach(ctx, ny, nz)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in), optional :: ny, nz
    kmm_ctx_handle = ctx
    if (present(ny)) kmm_nyp = ny + 1
    if (present(nz)) kmm_nz = nz
  end subroutine
  subroutine kmm_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= KMM_OK) then
      write(*,*) where, ': ', kmm_message(rc)
      stop
    end if
  end subroutine
end module kmm_dropin_state

module derivs
  use kmm_dropin_state
  implicit none
contains
  subroutine cderiv(f, df)
    complex(8) :: f(*), df(*)
    call kmm_check(kmm_cderiv(kmm_ctx_handle, f, df), 'cderiv')
  end subroutine
  subroutine gradscl(scalar, sclx, scly, sclz)
    complex(8) :: scalar(*), sclx(*), scly(*), sclz(*)
    call kmm_check(kmm_gradscl(kmm_ctx_handle, scalar, sclx, scly, sclz), 'gradscl')
  end subroutine
  subroutine derivscji(c11, c12, c13, c21, c22, c23, c31, c32, c33, &
                       dc111, dc112, dc113, dc211, dc212, dc213, dc311, dc312, dc313, &
                       dc121, dc122, dc123, dc221, dc222, dc223, dc321, dc322, dc323, &
                       dc131, dc132, dc133, dc231, dc232, dc233, dc331, dc332, dc333)
    complex(8), target :: c11(*), c12(*), c13(*), c21(*), c22(*), c23(*), c31(*), c32(*), c33(*)
    complex(8), target :: dc111(*), dc112(*), dc113(*), dc211(*), dc212(*), dc213(*), dc311(*), dc312(*), dc313(*)
    complex(8), target :: dc121(*), dc122(*), dc123(*), dc221(*), dc222(*), dc223(*), dc321(*), dc322(*), dc323(*)
    complex(8), target :: dc131(*), dc132(*), dc133(*), dc231(*), dc232(*), dc233(*), dc331(*), dc332(*), dc333(*)
    type(c_ptr) :: c(9), d(27)
    c = [c_loc(c11), c_loc(c12), c_loc(c13), c_loc(c21), c_loc(c22), c_loc(c23), c_loc(c31), c_loc(c32), &
         c_loc(c33)]
    d = [c_loc(dc111), c_loc(dc112), c_loc(dc113), c_loc(dc211), c_loc(dc212), c_loc(dc213), c_loc(dc311), &
         c_loc(dc312), c_loc(dc313), c_loc(dc121), c_loc(dc122), c_loc(dc123), c_loc(dc221), c_loc(dc222), &
         c_loc(dc223), c_loc(dc321), c_loc(dc322), c_loc(dc323), c_loc(dc131), c_loc(dc132), c_loc(dc133), &
         c_loc(dc231), c_loc(dc232), c_loc(dc233), c_loc(dc331), c_loc(dc332), c_loc(dc333)]
    call kmm_check(kmm_derivscji(kmm_ctx_handle, c, d), 'derivscji')
  end subroutine
end module derivs

module solvers
  use kmm_dropin_state
  implicit none
  interface vcw3dp
    module procedure vcw3dp_ns, vcw3dp_scalar, vcw3dp_scalar_polymer
  end interface
contains
  subroutine solve(u, x, g, dyde, bctop, bcbot, ib, a, wrkc)
    complex(8) :: u(*), bctop(*), bcbot(*), wrkc(*)
    real(8) :: x, g, dyde, a(*)
    integer :: ib
    call kmm_check(kmm_solve(kmm_ctx_handle, u, x, g, dyde, bctop, bcbot, int(ib, c_int)), 'solve')
  end subroutine
  subroutine penta(u, x, g, dyde, ib, at, bt, gtop, ab, bb, gbot, a, wrkc)
    complex(8) :: u(*), gtop(*), gbot(*), wrkc(*)
    real(8) :: x, g, dyde, at, bt, ab, bb, a(*)
    integer :: ib
    call kmm_check(kmm_penta(kmm_ctx_handle, u, x, g, dyde, int(ib, c_int), at, bt, gtop, ab, bb, gbot), 'penta')
  end subroutine
  subroutine phirhs(v, wrkc, wrk1, fn, fnm1, bcbot, bctop, var)
    complex(8) :: v(*), wrkc(*), wrk1(*), fn(*), fnm1(*), bcbot(*), bctop(*)
    real(8) :: var
    call kmm_check(kmm_phirhs(kmm_ctx_handle, v, fn, fnm1, bcbot, bctop, var), 'phirhs')
  end subroutine
  subroutine nonlin(fn, omz, gn, wrkc, wrk1)
    complex(8) :: fn(*), omz(*), gn(*), wrkc(*), wrk1(*)
    call kmm_check(kmm_nonlin(kmm_ctx_handle, fn, omz, gn), 'nonlin')
  end subroutine
  subroutine veloc(u, v, w, u0, w0, omy, wrkc)
    complex(8) :: u(*), v(*), w(*), omy(*), wrkc(*)
    real(8) :: u0(*), w0(*)
    call kmm_check(kmm_veloc(kmm_ctx_handle, u, v, w, u0, w0, omy), 'veloc')
  end subroutine
  subroutine vort(u, v, w, omx, omz)
    complex(8) :: u(*), v(*), w(*), omx(*), omz(*)
    call kmm_check(kmm_vort(kmm_ctx_handle, u, v, w, omx, omz), 'vort')
  end subroutine
  subroutine uzero(u0, h1, h1l, uflag)
    real(8) :: u0(*), h1(*), h1l(*)
    integer :: uflag
    call kmm_check(kmm_uzero(kmm_ctx_handle, u0, h1, h1l, int(uflag, c_int)), 'uzero')
  end subroutine
  ! vcw3dp, NS list (Lu, Lv, Lw feed the inertial particles the step tracks itself)
  subroutine vcw3dp_ns(u, v, w, omx, omy, omz, fn, gn, u11, u12, u13, u21, u22, u23, u31, u32, u33, Lu, Lv, Lw)
    complex(8) :: u(kmm_nyp, kmm_nz, *), v(kmm_nyp, kmm_nz, *), w(kmm_nyp, kmm_nz, *)
    complex(8) :: omx(kmm_nyp, kmm_nz, *), omy(kmm_nyp, kmm_nz, *), omz(kmm_nyp, kmm_nz, *)
    complex(8) :: fn(kmm_nyp, kmm_nz, *), gn(kmm_nyp, kmm_nz, *), Lu(kmm_nyp, kmm_nz, *)
    complex(8) :: Lv(kmm_nyp, kmm_nz, *), Lw(kmm_nyp, kmm_nz, *)
    complex(8), target :: u11(kmm_nyp, kmm_nz, *), u12(kmm_nyp, kmm_nz, *), u13(kmm_nyp, kmm_nz, *)
    complex(8), target :: u21(kmm_nyp, kmm_nz, *), u22(kmm_nyp, kmm_nz, *), u23(kmm_nyp, kmm_nz, *)
    complex(8), target :: u31(kmm_nyp, kmm_nz, *), u32(kmm_nyp, kmm_nz, *), u33(kmm_nyp, kmm_nz, *)
    type(c_ptr), target :: g(9)
    g = [c_loc(u11), c_loc(u12), c_loc(u13), c_loc(u21), c_loc(u22), c_loc(u23), c_loc(u31), c_loc(u32), &
         c_loc(u33)]
    call kmm_check(kmm_vcw3dp(kmm_ctx_handle, u, v, w, omx, omy, omz, fn, gn, c_loc(g), c_null_ptr, c_null_ptr, &
                              c_null_ptr, c_null_ptr, c_null_ptr, int(kmm_it, c_int)), 'vcw3dp')
  end subroutine
  ! vcw3dp, SCALAR list
  subroutine vcw3dp_scalar(u, v, w, omx, omy, omz, fn, gn, u11, u12, u13, u21, u22, u23, u31, u32, u33, &
                           scalar, sclx, scly, sclz, scn, Lu, Lv, Lw)
    complex(8) :: u(kmm_nyp, kmm_nz, *), v(kmm_nyp, kmm_nz, *), w(kmm_nyp, kmm_nz, *)
    complex(8) :: omx(kmm_nyp, kmm_nz, *), omy(kmm_nyp, kmm_nz, *), omz(kmm_nyp, kmm_nz, *)
    complex(8) :: fn(kmm_nyp, kmm_nz, *), gn(kmm_nyp, kmm_nz, *), Lu(kmm_nyp, kmm_nz, *)
    complex(8) :: Lv(kmm_nyp, kmm_nz, *), Lw(kmm_nyp, kmm_nz, *)
    complex(8), target :: u11(kmm_nyp, kmm_nz, *), u12(kmm_nyp, kmm_nz, *), u13(kmm_nyp, kmm_nz, *)
    complex(8), target :: u21(kmm_nyp, kmm_nz, *), u22(kmm_nyp, kmm_nz, *), u23(kmm_nyp, kmm_nz, *)
    complex(8), target :: u31(kmm_nyp, kmm_nz, *), u32(kmm_nyp, kmm_nz, *), u33(kmm_nyp, kmm_nz, *)
    complex(8), target :: scalar(kmm_nyp, kmm_nz, *), sclx(kmm_nyp, kmm_nz, *), scly(kmm_nyp, kmm_nz, *)
    complex(8), target :: sclz(kmm_nyp, kmm_nz, *), scn(kmm_nyp, kmm_nz, *)
    type(c_ptr), target :: g(9)
    g = [c_loc(u11), c_loc(u12), c_loc(u13), c_loc(u21), c_loc(u22), c_loc(u23), c_loc(u31), c_loc(u32), &
         c_loc(u33)]
    call kmm_check(kmm_vcw3dp(kmm_ctx_handle, u, v, w, omx, omy, omz, fn, gn, c_loc(g), c_loc(scalar), &
                              c_loc(sclx), c_loc(scly), c_loc(sclz), c_loc(scn), int(kmm_it, c_int)), 'vcw3dp')
  end subroutine
  ! vcw3dp, SCALAR + POLYMER list (the reference's polymer build)
  subroutine vcw3dp_scalar_polymer(u, v, w, omx, omy, omz, fn, gn, u11, u12, u13, u21, u22, u23, u31, u32, u33, &
                                   scalar, sclx, scly, sclz, scn, &
                                   c11, c12, c13, c21, c22, c23, c31, c32, c33, &
                                   dc111, dc112, dc113, dc211, dc212, dc213, dc311, dc312, dc313, &
                                   dc121, dc122, dc123, dc221, dc222, dc223, dc321, dc322, dc323, &
                                   dc131, dc132, dc133, dc231, dc232, dc233, dc331, dc332, dc333, &
                                   c11n, c12n, c13n, c22n, c23n, c33n, &
                                   str11n, str12n, str13n, str22n, str23n, str33n, &
                                   qp11, qp12, qp13, qp22, qp23, qp33, Lu, Lv, Lw)
    complex(8) :: u(kmm_nyp, kmm_nz, *), v(kmm_nyp, kmm_nz, *), w(kmm_nyp, kmm_nz, *)
    complex(8) :: omx(kmm_nyp, kmm_nz, *), omy(kmm_nyp, kmm_nz, *), omz(kmm_nyp, kmm_nz, *)
    complex(8) :: fn(kmm_nyp, kmm_nz, *), gn(kmm_nyp, kmm_nz, *), Lu(kmm_nyp, kmm_nz, *)
    complex(8) :: Lv(kmm_nyp, kmm_nz, *), Lw(kmm_nyp, kmm_nz, *)
    complex(8), target :: u11(kmm_nyp, kmm_nz, *), u12(kmm_nyp, kmm_nz, *), u13(kmm_nyp, kmm_nz, *)
    complex(8), target :: u21(kmm_nyp, kmm_nz, *), u22(kmm_nyp, kmm_nz, *), u23(kmm_nyp, kmm_nz, *)
    complex(8), target :: u31(kmm_nyp, kmm_nz, *), u32(kmm_nyp, kmm_nz, *), u33(kmm_nyp, kmm_nz, *)
    complex(8), target :: scalar(kmm_nyp, kmm_nz, *), sclx(kmm_nyp, kmm_nz, *), scly(kmm_nyp, kmm_nz, *)
    complex(8), target :: sclz(kmm_nyp, kmm_nz, *), scn(kmm_nyp, kmm_nz, *)
    complex(8), target :: c11(kmm_nyp, kmm_nz, *), c12(kmm_nyp, kmm_nz, *), c13(kmm_nyp, kmm_nz, *)
    complex(8), target :: c21(kmm_nyp, kmm_nz, *), c22(kmm_nyp, kmm_nz, *), c23(kmm_nyp, kmm_nz, *)
    complex(8), target :: c31(kmm_nyp, kmm_nz, *), c32(kmm_nyp, kmm_nz, *), c33(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc111(kmm_nyp, kmm_nz, *), dc112(kmm_nyp, kmm_nz, *), dc113(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc211(kmm_nyp, kmm_nz, *), dc212(kmm_nyp, kmm_nz, *), dc213(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc311(kmm_nyp, kmm_nz, *), dc312(kmm_nyp, kmm_nz, *), dc313(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc121(kmm_nyp, kmm_nz, *), dc122(kmm_nyp, kmm_nz, *), dc123(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc221(kmm_nyp, kmm_nz, *), dc222(kmm_nyp, kmm_nz, *), dc223(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc321(kmm_nyp, kmm_nz, *), dc322(kmm_nyp, kmm_nz, *), dc323(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc131(kmm_nyp, kmm_nz, *), dc132(kmm_nyp, kmm_nz, *), dc133(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc231(kmm_nyp, kmm_nz, *), dc232(kmm_nyp, kmm_nz, *), dc233(kmm_nyp, kmm_nz, *)
    complex(8), target :: dc331(kmm_nyp, kmm_nz, *), dc332(kmm_nyp, kmm_nz, *), dc333(kmm_nyp, kmm_nz, *)
    complex(8), target :: c11n(kmm_nyp, kmm_nz, *), c12n(kmm_nyp, kmm_nz, *), c13n(kmm_nyp, kmm_nz, *)
    complex(8), target :: c22n(kmm_nyp, kmm_nz, *), c23n(kmm_nyp, kmm_nz, *), c33n(kmm_nyp, kmm_nz, *)
    complex(8), target :: str11n(kmm_nyp, kmm_nz, *), str12n(kmm_nyp, kmm_nz, *), str13n(kmm_nyp, kmm_nz, *)
    complex(8), target :: str22n(kmm_nyp, kmm_nz, *), str23n(kmm_nyp, kmm_nz, *), str33n(kmm_nyp, kmm_nz, *)
    complex(8), target :: qp11(kmm_nyp, kmm_nz, *), qp12(kmm_nyp, kmm_nz, *), qp13(kmm_nyp, kmm_nz, *)
    complex(8), target :: qp22(kmm_nyp, kmm_nz, *), qp23(kmm_nyp, kmm_nz, *), qp33(kmm_nyp, kmm_nz, *)
    type(c_ptr) :: g(9), c(9), d(27), cn(6), sn(6), q(6)
    g = [c_loc(u11), c_loc(u12), c_loc(u13), c_loc(u21), c_loc(u22), c_loc(u23), c_loc(u31), c_loc(u32), &
         c_loc(u33)]
    c = [c_loc(c11), c_loc(c12), c_loc(c13), c_loc(c21), c_loc(c22), c_loc(c23), c_loc(c31), c_loc(c32), &
         c_loc(c33)]
    d = [c_loc(dc111), c_loc(dc112), c_loc(dc113), c_loc(dc211), c_loc(dc212), c_loc(dc213), c_loc(dc311), &
         c_loc(dc312), c_loc(dc313), c_loc(dc121), c_loc(dc122), c_loc(dc123), c_loc(dc221), c_loc(dc222), &
         c_loc(dc223), c_loc(dc321), c_loc(dc322), c_loc(dc323), c_loc(dc131), c_loc(dc132), c_loc(dc133), &
         c_loc(dc231), c_loc(dc232), c_loc(dc233), c_loc(dc331), c_loc(dc332), c_loc(dc333)]
    cn = [c_loc(c11n), c_loc(c12n), c_loc(c13n), c_loc(c22n), c_loc(c23n), c_loc(c33n)]
    sn = [c_loc(str11n), c_loc(str12n), c_loc(str13n), c_loc(str22n), c_loc(str23n), c_loc(str33n)]
    q = [c_loc(qp11), c_loc(qp12), c_loc(qp13), c_loc(qp22), c_loc(qp23), c_loc(qp33)]
    call kmm_check(kmm_vcw3dp_poly(kmm_ctx_handle, u, v, w, omx, omy, omz, fn, gn, g, c_loc(scalar), &
                                   c_loc(sclx), c_loc(scly), c_loc(sclz), c_loc(scn), c, d, cn, sn, q, &
                                   int(kmm_it, c_int)), 'vcw3dp')
  end subroutine
end module solvers

module helpers
  use kmm_dropin_state
  implicit none
contains
  subroutine norm(a)
    complex(8) :: a(*)
    call kmm_check(kmm_norm(kmm_ctx_handle, a), 'norm')
  end subroutine
  subroutine polyforce(s11, s12, s13, s22, s23, s33, t1, t2, t3)
    complex(8) :: s11(*), s12(*), s13(*), s22(*), s23(*), s33(*), t1(*), t2(*), t3(*)
    call kmm_check(kmm_polyforce(kmm_ctx_handle, s11, s12, s13, s22, s23, s33, t1, t2, t3), 'polyforce')
  end subroutine
  subroutine subforce(gn, fn, omz, t1, t2, t3)
    complex(8) :: gn(*), fn(*), omz(*), t1(*), t2(*), t3(*)
    call kmm_check(kmm_subforce(kmm_ctx_handle, gn, fn, omz, t1, t2, t3), 'subforce')
  end subroutine
  subroutine polyNL(c11NL, c12NL, c13NL, c22NL, c23NL, c33NL, &
                    c11n, c12n, c13n, c22n, c23n, c33n, &
                    c11nm1, c12nm1, c13nm1, c22nm1, c23nm1, c33nm1, &
                    str11n, str12n, str13n, str22n, str23n, str33n, &
                    str11nm1, str12nm1, str13nm1, str22nm1, str23nm1, str33nm1)
    complex(8), target :: c11NL(*), c12NL(*), c13NL(*), c22NL(*), c23NL(*), c33NL(*)
    complex(8), target :: c11n(*), c12n(*), c13n(*), c22n(*), c23n(*), c33n(*)
    complex(8), target :: c11nm1(*), c12nm1(*), c13nm1(*), c22nm1(*), c23nm1(*), c33nm1(*)
    complex(8), target :: str11n(*), str12n(*), str13n(*), str22n(*), str23n(*), str33n(*)
    complex(8), target :: str11nm1(*), str12nm1(*), str13nm1(*), str22nm1(*), str23nm1(*), str33nm1(*)
    type(c_ptr) :: o(6), a(6), b(6), s(6), t(6)
    o = [c_loc(c11NL), c_loc(c12NL), c_loc(c13NL), c_loc(c22NL), c_loc(c23NL), c_loc(c33NL)]
    a = [c_loc(c11n), c_loc(c12n), c_loc(c13n), c_loc(c22n), c_loc(c23n), c_loc(c33n)]
    b = [c_loc(c11nm1), c_loc(c12nm1), c_loc(c13nm1), c_loc(c22nm1), c_loc(c23nm1), c_loc(c33nm1)]
    s = [c_loc(str11n), c_loc(str12n), c_loc(str13n), c_loc(str22n), c_loc(str23n), c_loc(str33n)]
    t = [c_loc(str11nm1), c_loc(str12nm1), c_loc(str13nm1), c_loc(str22nm1), c_loc(str23nm1), c_loc(str33nm1)]
    call kmm_check(kmm_polyNL(kmm_ctx_handle, o, a, b, s, t), 'polyNL')
  end subroutine
end module helpers
