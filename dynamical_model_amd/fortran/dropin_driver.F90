! dropin_driver.F90 -- the reference's program flow with the MI355X mod_moloch.
!
! Follows nompi/regcm_dynamical_core.F90: parameters and geometry (`param`,
! :93-194), fill_atmosphere (mod_atmosphere.F90:171-235, the reference's own
! code), init_moloch, then moloch(i, mod(i,iprint)==0) for i = 1..nstep
! (:49-53).  The terrain is one hill h*exp(-ln2*(r/hw)**2) so the same state
! can be built by the package's case builder.  Instead of the netCDF output
! it writes the final state to a raw file for the parity test.
!
! Usage: dropin_driver jx iy kz nstep hill_m hill_halfwidth precision(0|1) out.bin
program dropin_driver
  use mod_intkinds
  use mod_realkinds
  use mod_constants
  use mod_dynparam
  use mod_runparams
  use mod_memutil
  use mod_atmosphere
  use mod_zita
  use mod_moloch
  implicit none
  character(len=256) :: arg , outfile
  real(rkx) :: hill , hw , rr , dl
  integer(ik4) :: i , j , k , u , nsteps , prec
  integer(ik4) , parameter :: iout_drv = 2
  logical :: lprint

  call get_command_argument(1,arg) ; read(arg,*) jx
  call get_command_argument(2,arg) ; read(arg,*) iy
  call get_command_argument(3,arg) ; read(arg,*) kz
  call get_command_argument(4,arg) ; read(arg,*) nsteps
  call get_command_argument(5,arg) ; read(arg,*) hill
  call get_command_argument(6,arg) ; read(arg,*) hw
  call get_command_argument(7,arg) ; read(arg,*) prec
  call get_command_argument(8,outfile)
  iprint = 2

  call memory_init
  call initparam
  global_dot_jstart = 1
  global_dot_istart = 1
  global_dot_jend = jx
  global_dot_iend = iy
  global_cross_jstart = 1
  global_cross_istart = 1
  global_cross_jend = jx
  global_cross_iend = iy
  call setup_model_indexes

  ! param (nompi/regcm_dynamical_core.F90:93-194)
  ds = 3.0
  dx = ds * 1000.0_rkx
  rdx = 1.0_rkx/dx
  dt = 30.0_rkx
  rdt = 1.0_rkx/dt
  mo_ztop = 36000_rkx
  mo_h = 12000_rkx
  mo_a0 = 0.0_rkx
  mo_nadv = 3
  mo_nsound = 5
  mo_nzfilt = kz/5
  mo_anu2 = 0.6_rkx
  nqx = 5
  iqfrst = iqc
  iqlst = iqs
  call allocate_mod_runparams
  call allocate_mod_atm_interface
  call allocate_moloch
  call model_zitaf(zita,mo_ztop)
  call model_zitah(zitah,mo_ztop)
  mo_dzita = zita(kz)

  mddom%ht = 0.0_rkx
  do i = ide1 , ide2
    do j = jde1 , jde2
      rr = sqrt((real(j,rkx)-0.5_rkx*real(jx+1,rkx))**2 + &
                (real(i,rkx)-0.5_rkx*real(iy+1,rkx))**2)
      if ( hw > 0.0_rkx ) then
        mddom%ht(j,i) = egrav*hill*exp(-log(2.0_rkx)*(rr/hw)**2)
      end if
    end do
  end do
  mddom%ht(jde1-1,:) = mddom%ht(jde2,:)
  mddom%ht(jde2+1,:) = mddom%ht(jde1,:)
  mddom%ht(:,ide1-1) = mddom%ht(:,ide2)
  mddom%ht(:,ide2+1) = mddom%ht(:,ide1)
  mddom%msfx = 1.0_rkx
  mddom%msfu = 1.0_rkx
  mddom%msfv = 1.0_rkx
  dl = raddeg * dx/earthrad
  do i = ide1 , ide2
    do j = jde1 , jde2
      mddom%xlat(j,i) = dl * (i-real(iy,rkx)*d_half-0.5_rkx)
      mddom%ulat(j,i) = mddom%xlat(j,i)
      mddom%vlat(j,i) = dl * (i-real(iy,rkx)*d_half)
    end do
  end do
  do i = ice1 , ice2
    do j = jdi1 , jdi2
      mddom%hx(j,i) = (mddom%ht(j,i) - mddom%ht(j-1,i)) * &
                       mddom%msfu(j,i) * rdx * regrav
    end do
  end do
  do i = idi1 , idi2
    do j = jce1 , jce2
      mddom%hy(j,i) = (mddom%ht(j,i) - mddom%ht(j,i-1)) * rdx * regrav
    end do
  end do
  mddom%hx(jde1ga,ice1:ice2) = mddom%hx(jde2,ice1:ice2)
  mddom%hx(jde2ga,ice1:ice2) = mddom%hx(jde1,ice1:ice2)
  mddom%hy(jce1:jce2,ide1ga) = mddom%hy(jce1:jce2,ide2)
  mddom%hy(jce1:jce2,ide2ga) = mddom%hy(jce1:jce2,ide1)
  do k = 1 , kz
    do i = ice1 , ice2
      do j = jce1 , jce2
        mo_atm%zeta(j,i,k) = md_zeta(zitah(k),mddom%ht(j,i),mo_ztop,mo_h,mo_a0)
        mo_atm%fmz(j,i,k) = md_fmz(zitah(k),mddom%ht(j,i),mo_ztop,mo_h,mo_a0)
      end do
    end do
  end do
  mo_atm%fmz(jce1ga,ice1:ice2,:) = mo_atm%fmz(jce2,ice1:ice2,:)
  mo_atm%fmz(jce2ga,ice1:ice2,:) = mo_atm%fmz(jce1,ice1:ice2,:)
  mo_atm%fmz(jce1:jce2,ice1ga,:) = mo_atm%fmz(jce1:jce2,ice2,:)
  mo_atm%fmz(jce1:jce2,ice2ga,:) = mo_atm%fmz(jce1:jce2,ice1,:)
  do k = 1 , kzp1
    do i = ice1 , ice2
      do j = jce1 , jce2
        mo_atm%fmzf(j,i,k) = md_fmz(zita(k),mddom%ht(j,i),mo_ztop,mo_h,mo_a0)
      end do
    end do
  end do
  mddom%msfu(jde1ga,ide1:ide2) = mddom%msfu(jde2,ide1:ide2)
  mddom%msfu(jde2ga,ide1:ide2) = mddom%msfu(jde1,ide1:ide2)
  mddom%msfv(jde1:jde2,ide1ga) = mddom%msfv(jde1:jde2,ide2)
  mddom%msfv(jde1:jde2,ide2ga) = mddom%msfv(jde1:jde2,ide1)

  ! init (:196-201)
  moloch_precision = prec
  call fill_atmosphere
  call init_moloch

  ! output(istep) before moloch(istep) (nompi/regcm_dynamical_core.F90:48-52,
  ! :202-215): ps.nc and t.nc beside the state file, every iout_drv steps
  k = index(outfile,'/',back=.true.)
  if ( k > 0 ) then
    call moloch_output_open(outfile(1:k-1))
  else
    call moloch_output_open('.')
  end if
  do i = 1 , nsteps
    if ( i == 1 .or. mod(i,iout_drv) == 0 ) call moloch_output_write
    lprint = mod(i,iprint) == 0
    call moloch(i,lprint)
  end do
  call moloch_output_close

  open(newunit=u,file=trim(outfile),access='stream',form='unformatted', &
       status='replace')
  write(u) int(jx,4) , int(iy,4) , int(kz,4)
  write(u) mo_atm%u(1:jx,1:iy,1:kz) , mo_atm%v(1:jx,1:iy,1:kz) , &
           mo_atm%w(1:jx,1:iy,1:kzp1) , mo_atm%pai(1:jx,1:iy,1:kz) , &
           mo_atm%tetav(1:jx,1:iy,1:kz) , mo_atm%qx(1:jx,1:iy,1:kz,1) , &
           mo_atm%t(1:jx,1:iy,1:kz) , sfs%ps(1:jx,1:iy)
  close(u)
  call moloch_finalize
end program dropin_driver
