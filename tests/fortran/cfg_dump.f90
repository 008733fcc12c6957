! Prints every field of kmm_config as read by kmm_config_read, one
! "name value" pair per line (integers as integers, reals with 17 digits).
!   cfg_dump <dns.config>
program cfg_dump
  use kmm_c
  implicit none
  type(kmm_config) :: c
  character(len=1024) :: path
  integer(c_int) :: rc
  call get_command_argument(1, path)
  rc = kmm_config_read(trim(path)//c_null_char, c)
  write(*, '(a, i0)') 'rc ', rc
  call i('nx', c%nx); call i('ny', c%ny); call i('nz', c%nz); call i('nsteps', c%nsteps)
  call i('iprnfrq', c%iprnfrq); call r('dt', c%dt); call r('gain', c%gain); call r('ugain', c%ugain)
  call r('theta', c%theta); call r('yl', c%yl); call r('zl', c%zl); call r('xl', c%xl); call r('re', c%re)
  call r('uinf', c%uinf); call i('kwall', c%kwall); call i('perturbtime', c%perturbtime)
  call r('r_tau', c%r_tau); call i('irstrt', c%irstrt); call i('crstrt', c%crstrt)
  call i('flow_select', c%flow_select); call i('scalar', c%scalar); call i('scl_flag', c%scl_flag)
  call r('diff', c%diff); call r('deltaT', c%deltaT); call i('src_start', c%src_start)
  call i('src_stop', c%src_stop); call i('cadence', c%cadence); call i('precision', c%precision)
  call i('particle_flag', c%particle_flag); call r('ratio', c%ratio); call r('ap', c%ap)
  call r('c_mu', c%c_mu); call i('cd_switch', c%cd_switch); call i('polymer', c%polymer)
  call i('itarget', c%itarget); call i('ipeter', c%ipeter); call i('ipolyflag', c%ipolyflag)
  call r('zlmax', c%zlmax); call r('tpoly', c%tpoly); call r('alpha_poly', c%alpha_poly)
  call r('diffpoly', c%diffpoly); call r('c11z', c%c11z); call r('c22z', c%c22z); call r('c33z', c%c33z)
  call r('qbeta', c%qbeta); call i('cmf', c%cmf); call r('ubulk', c%ubulk)
  call i('particle_swirl', c%particle_swirl); call r('sigmax', c%sigmax); call r('sigmay', c%sigmay)
  call r('sigmaz', c%sigmaz); call r('gravity', c%gravity); call i('bftail', c%bftail)
  call r('xstart', c%xstart)
contains
  subroutine i(name, v)
    character(len=*), intent(in) :: name
    integer(c_int), intent(in) :: v
    write(*, '(a, 1x, i0)') name, v
  end subroutine
  subroutine r(name, v)
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: v
    write(*, '(a, 1x, es25.17)') name, v
  end subroutine
end program cfg_dump
