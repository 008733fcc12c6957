! kmm.f90 -- Fortran binding of the C ABI in kmm.h (iso_c_binding).
!
! Module kmm_c mirrors struct kmm_config field by field and declares the
! entry points a Fortran host of the reference's time-step path calls.
! Arrays are passed in the reference's own layout: spectral fields
! complex(8) (nyp,nz,nxh), mean profiles real(8) (nyp), IB integrals and
! the geometry (nyp,mz,mx).  Strings are NUL-terminated (append c_null_char).
module kmm_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: KMM_OK = 0, KMM_E_CFL = 1, KMM_E_PARTICLE = 2, KMM_E_HIP = 3, &
                               KMM_E_CONFIG = 4, KMM_E_IO = 5, KMM_E_ARG = 6
  ! field ids of kmm_put_field / kmm_get_field (kmm.h)
  integer(c_int), parameter :: KMM_V = 0, KMM_FN = 1, KMM_FNM1 = 2, KMM_OMY = 3, KMM_GN = 4, KMM_GNM1 = 5, &
                               KMM_U0 = 6, KMM_H1N = 7, KMM_H1NM1 = 8, KMM_W0 = 9, KMM_H3N = 10, &
                               KMM_H3NM1 = 11, KMM_FXINTG = 12, KMM_FYINTG = 13, KMM_FZINTG = 14, &
                               KMM_SCALAR = 15, KMM_SCN = 16, KMM_SCNM1 = 17, KMM_C11 = 18

  type, bind(C) :: kmm_config
    integer(c_int) :: nx, ny, nz, nsteps, iprnfrq
    real(c_double) :: dt, gain, ugain, theta, yl, zl, xl, re, uinf
    integer(c_int) :: kwall, perturbtime
    real(c_double) :: r_tau
    integer(c_int) :: irstrt, crstrt, flow_select, scalar, scl_flag
    real(c_double) :: diff, deltaT
    integer(c_int) :: src_start, src_stop, cadence, precision
    integer(c_int) :: particle_flag
    real(c_double) :: ratio, ap, c_mu
    integer(c_int) :: cd_switch
    integer(c_int) :: polymer, itarget, ipeter, ipolyflag
    real(c_double) :: zlmax, tpoly, alpha_poly, diffpoly, c11z, c22z, c33z, qbeta
    integer(c_int) :: cmf
    real(c_double) :: ubulk
    integer(c_int) :: particle_swirl
    real(c_double) :: sigmax, sigmay, sigmaz
    real(c_double) :: gravity
    integer(c_int) :: bftail
    real(c_double) :: xstart
  end type

  interface
    type(c_ptr) function kmm_status_string(status) bind(C)
      import; integer(c_int), value :: status
    end function
    integer(c_int) function kmm_config_read(path, cfg) bind(C)
      import; character(kind=c_char) :: path(*); type(kmm_config) :: cfg
    end function
    integer(c_int) function kmm_create(cfg, device, ctx) bind(C)
      import; type(kmm_config) :: cfg; integer(c_int), value :: device; type(c_ptr) :: ctx
    end function
    integer(c_int) function kmm_destroy(ctx) bind(C)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function kmm_set_geometry(ctx, imatrix) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int32_t) :: imatrix(*)
    end function
    integer(c_int) function kmm_put_field(ctx, id, a) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: id; real(c_double) :: a(*)
    end function
    integer(c_int) function kmm_get_field(ctx, id, a) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: id; real(c_double) :: a(*)
    end function
    integer(c_int) function kmm_startup(ctx) bind(C)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function kmm_step(ctx, nsteps) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: nsteps
    end function
    integer(c_int) function kmm_last_cfl(ctx, cflmax, it) bind(C)
      import; type(c_ptr), value :: ctx; real(c_double) :: cflmax; integer(c_int) :: it
    end function
    integer(c_int) function kmm_read_restart(ctx, path) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: path(*)
    end function
    integer(c_int) function kmm_write_restart(ctx, path) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: path(*)
    end function
    integer(c_int) function kmm_read_crestart(ctx, path) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: path(*)
    end function
    integer(c_int) function kmm_write_crestart(ctx, path) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: path(*)
    end function
    integer(c_int) function kmm_read_particles(ctx, path) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: path(*)
    end function
    integer(c_int) function kmm_write_particles(ctx, path, it) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: path(*); integer(c_int), value :: it
    end function
    integer(c_int) function kmm_set_graph(ctx, steps) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: steps
    end function
    integer(c_int) function kmm_write_outputs(ctx, dir, first) bind(C)
      import; type(c_ptr), value :: ctx; character(kind=c_char) :: dir(*); integer(c_int), value :: first
    end function
    ! per-routine entries (reference argument meaning; work arrays internal)
    integer(c_int) function kmm_cderiv(ctx, f, df) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: f(*), df(*)
    end function
    integer(c_int) function kmm_solve(ctx, u, x, g, dyde, bctop, bcbot, ib) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: u(*), bctop(*), bcbot(*)
      real(c_double), value :: x, g, dyde; integer(c_int), value :: ib
    end function
    integer(c_int) function kmm_penta(ctx, u, x, g, dyde, ib, at, bt, gtop, ab, bb, gbot) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: u(*), gtop(*), gbot(*)
      real(c_double), value :: x, g, dyde, at, bt, ab, bb; integer(c_int), value :: ib
    end function
    integer(c_int) function kmm_phirhs(ctx, v, fn, fnm1, bcbot, bctop, var) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: v(*), fn(*), fnm1(*), bcbot(*), bctop(*)
      real(c_double), value :: var
    end function
    integer(c_int) function kmm_nonlin(ctx, fn, omz, gn) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: fn(*), omz(*), gn(*)
    end function
    integer(c_int) function kmm_veloc(ctx, u, v, w, u0, w0, omy) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: u(*), v(*), w(*), omy(*)
      real(c_double) :: u0(*), w0(*)
    end function
    integer(c_int) function kmm_vort(ctx, u, v, w, omx, omz) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: u(*), v(*), w(*), omx(*), omz(*)
    end function
    integer(c_int) function kmm_gradscl(ctx, scalar, sclx, scly, sclz) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: scalar(*), sclx(*), scly(*), sclz(*)
    end function
    integer(c_int) function kmm_derivscji(ctx, cij, dc) bind(C)
      import; type(c_ptr), value :: ctx; type(c_ptr) :: cij(9), dc(27)
    end function
    integer(c_int) function kmm_polyforce(ctx, s11, s12, s13, s22, s23, s33, t1, t2, t3) bind(C)
      import; type(c_ptr), value :: ctx
      complex(c_double_complex) :: s11(*), s12(*), s13(*), s22(*), s23(*), s33(*), t1(*), t2(*), t3(*)
    end function
    integer(c_int) function kmm_subforce(ctx, gn, fn, omz, t1, t2, t3) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: gn(*), fn(*), omz(*), t1(*), t2(*), t3(*)
    end function
    integer(c_int) function kmm_polyNL(ctx, cNL, cn, cnm1, strn, strnm1) bind(C, name="kmm_polyNL")
      import; type(c_ptr), value :: ctx; type(c_ptr) :: cNL(6), cn(6), cnm1(6), strn(6), strnm1(6)
    end function
    integer(c_int) function kmm_norm(ctx, a) bind(C)
      import; type(c_ptr), value :: ctx; complex(c_double_complex) :: a(*)
    end function
    integer(c_int) function kmm_uzero(ctx, u0, h1, h1l, uflag) bind(C)
      import; type(c_ptr), value :: ctx; real(c_double) :: u0(*), h1(*), h1l(*); integer(c_int), value :: uflag
    end function
    ! vcw3dp: uij is the address of 9 c_ptr (u11 .. u33) or c_null_ptr; the
    ! scalar arguments may be c_null_ptr (no SCALAR list)
    integer(c_int) function kmm_vcw3dp(ctx, u, v, w, omx, omy, omz, fn, gn, uij, scalar, sclx, scly, sclz, &
                                       scn, it) bind(C)
      import; type(c_ptr), value :: ctx
      complex(c_double_complex) :: u(*), v(*), w(*), omx(*), omy(*), omz(*), fn(*), gn(*)
      type(c_ptr), value :: uij, scalar, sclx, scly, sclz, scn
      integer(c_int), value :: it
    end function
    ! vcw3dp with the POLYMER list: cij(9), dc(27) in, cn(6), strn(6), qp(6) out
    integer(c_int) function kmm_vcw3dp_poly(ctx, u, v, w, omx, omy, omz, fn, gn, uij, scalar, sclx, scly, sclz, &
                                            scn, cij, dc, cn, strn, qp, it) bind(C)
      import; type(c_ptr), value :: ctx
      complex(c_double_complex) :: u(*), v(*), w(*), omx(*), omy(*), omz(*), fn(*), gn(*)
      type(c_ptr) :: uij(9)
      type(c_ptr), value :: scalar, sclx, scly, sclz, scn
      type(c_ptr) :: cij(9), dc(27), cn(6), strn(6), qp(6)
      integer(c_int), value :: it
    end function
  end interface

contains

  ! kmm_status_string as a Fortran string
  function kmm_message(status) result(msg)
    integer(c_int), intent(in) :: status
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(p) bind(C, name="strlen")
        import; type(c_ptr), value :: p
      end function
    end interface
    type(c_ptr) :: p
    p = kmm_status_string(status)
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

end module kmm_c
