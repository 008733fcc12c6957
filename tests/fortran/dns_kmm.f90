! Host program of the drop-in test: the reference's run sequence for a
! restart run (main_dns.f90: setstuff read, initial from setup/restart, the
! time loop, the last-restart write) through the C ABI.
!   dns_kmm <dns.config> <restart in> <restart out>
! The step count is NSTEPS of the configuration.  Exit status 0 on success;
! a failing call prints the library's message and stops with status 1.
program dns_kmm
  use kmm_c
  implicit none
  type(kmm_config) :: cfg
  type(c_ptr) :: ctx
  character(len=1024) :: cfgp, rin, rout
  call get_command_argument(1, cfgp)
  call get_command_argument(2, rin)
  call get_command_argument(3, rout)
  call chk(kmm_config_read(trim(cfgp)//c_null_char, cfg), 'config')
  call chk(kmm_create(cfg, 0_c_int, ctx), 'create')
  call chk(kmm_read_restart(ctx, trim(rin)//c_null_char), 'read restart')
  call chk(kmm_step(ctx, int(cfg%nsteps, c_int)), 'step')
  call chk(kmm_write_restart(ctx, trim(rout)//c_null_char), 'write restart')
  call chk(kmm_destroy(ctx), 'destroy')
contains
  subroutine chk(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= KMM_OK) then
      write(*,*) what, ': ', kmm_message(rc)
      stop 1
    end if
  end subroutine
end program dns_kmm
