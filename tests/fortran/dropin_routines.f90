! Host program of the drop-in modules' runtime test: the reference's
! procedure names from modules derivs/solvers/helpers (include/kmm_dropin.f90)
! on fields read from a raw file, results written to a raw file.
!   dropin_routines <dns.config> <fields in> <results out>
! in:  u, v, w, omy (complex(8) (nyp,nz,nxh) each, stream access)
! out: cderiv(u), vort's omx and omz, then vcw3dp's (NS list) fn, gn and
!      omz (in/out), norm(fn) -- complex(8) (nyp,nz,nxh) each
program dropin_routines
  use kmm_c
  use kmm_dropin_state
  use derivs, only: cderiv
  use solvers, only: vort, vcw3dp
  use helpers, only: norm
  implicit none
  type(kmm_config) :: cfg
  type(c_ptr) :: ctx
  character(len=1024) :: cfgp, fin, fout
  complex(8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), omy(:,:,:), du(:,:,:), omx(:,:,:), omz(:,:,:)
  complex(8), allocatable :: oz(:,:,:), fn(:,:,:), gn(:,:,:), g(:,:,:), lw(:,:,:)
  integer :: nyp, nz, nxh
  call get_command_argument(1, cfgp)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  call kmm_check(kmm_config_read(trim(cfgp)//c_null_char, cfg), 'config')
  call kmm_check(kmm_create(cfg, 0_c_int, ctx), 'create')
  call kmm_attach(ctx, cfg%ny, cfg%nz)
  nyp = cfg%ny + 1
  nz = cfg%nz
  nxh = cfg%nx / 2
  allocate(u(nyp,nz,nxh), v(nyp,nz,nxh), w(nyp,nz,nxh), omy(nyp,nz,nxh), du(nyp,nz,nxh), omx(nyp,nz,nxh), &
           omz(nyp,nz,nxh), oz(nyp,nz,nxh), fn(nyp,nz,nxh), gn(nyp,nz,nxh), g(nyp,nz,nxh), lw(nyp,nz,nxh))
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) u, v, w, omy
  close(10)
  g = (0.0d0, 0.0d0)
  lw = (0.0d0, 0.0d0)
  call cderiv(u, du)
  call vort(u, v, w, omx, omz)
  oz = omz
  kmm_it = 1
  call vcw3dp(u, v, w, omx, omy, oz, fn, gn, g, g, g, g, g, g, g, g, g, lw, lw, lw)
  call norm(fn)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) du, omx, omz, fn, gn, oz
  close(11)
  call kmm_check(kmm_destroy(ctx), 'destroy')
end program dropin_routines
