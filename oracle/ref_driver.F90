! ref_driver.F90 -- TEST INFRASTRUCTURE ONLY.
!
! A small driver of this project's own that links the reference modules which
! compile from their own sources (mod_constants, mod_zita, mod_atmosphere and
! their dependencies, built by `make -C oracle ref` into oracle/_ref) and
! writes what they compute to a raw binary file.  tests/golden/make_ref_vectors.py
! turns that file into the committed golden vectors that pin the CPU oracle
! and the product's case builder.
!
! The geometry set-up restates the loops of `param`
! (nompi/regcm_dynamical_core.F90:93-194); the physics (zita levels, md_zeta,
! md_fmz, gzita, the hydrostatic fill) is the reference's own code.
!
! Usage: ref_driver jx iy kz ztop zh a0 hill_m hill_halfwidth_cells out.bin
program ref_driver
  use mod_intkinds
  use mod_realkinds
  use mod_constants
  use mod_dynparam
  use mod_runparams
  use mod_memutil
  use mod_atmosphere
  use mod_zita
  implicit none
  character(len=256) :: arg , outfile
  real(rkx) :: hill , hw , rr , dl
  integer(ik4) :: i , j , k , u

  call get_command_argument(1,arg) ; read(arg,*) jx
  call get_command_argument(2,arg) ; read(arg,*) iy
  call get_command_argument(3,arg) ; read(arg,*) kz
  call get_command_argument(4,arg) ; read(arg,*) mo_ztop
  call get_command_argument(5,arg) ; read(arg,*) mo_h
  call get_command_argument(6,arg) ; read(arg,*) mo_a0
  call get_command_argument(7,arg) ; read(arg,*) hill
  call get_command_argument(8,arg) ; read(arg,*) hw
  call get_command_argument(9,outfile)

  call memory_init
  call initparam
  ! the serial decomposition of nompi/mod_mppparam.F90:109-119
  global_dot_jstart = 1
  global_dot_istart = 1
  global_dot_jend = jx
  global_dot_iend = iy
  global_cross_jstart = 1
  global_cross_istart = 1
  global_cross_jend = jx
  global_cross_iend = iy
  call setup_model_indexes

  dx = 3000.0_rkx
  rdx = 1.0_rkx/dx
  dt = 30.0_rkx
  mo_nadv = 3
  mo_nsound = 5
  nqx = 5
  iqfrst = iqc
  iqlst = iqs
  call allocate_mod_runparams
  call allocate_mod_atm_interface
  call model_zitaf(zita,mo_ztop)
  call model_zitah(zitah,mo_ztop)
  mo_dzita = zita(kz)

  ! terrain: a centred hill h*exp(-ln2*(r/hw)**2) as geopotential, with the
  ! periodic halo filled before the slopes are taken
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
  do k = 1 , kz
    do i = ice1 , ice2
      do j = jce1 , jce2
        mo_atm%zeta(j,i,k) = md_zeta(zitah(k),mddom%ht(j,i),mo_ztop,mo_h,mo_a0)
        mo_atm%fmz(j,i,k) = md_fmz(zitah(k),mddom%ht(j,i),mo_ztop,mo_h,mo_a0)
      end do
    end do
  end do
  do k = 1 , kzp1
    do i = ice1 , ice2
      do j = jce1 , jce2
        mo_atm%fmzf(j,i,k) = md_fmz(zita(k),mddom%ht(j,i),mo_ztop,mo_h,mo_a0)
      end do
    end do
  end do

  call fill_atmosphere

  open(newunit=u,file=trim(outfile),access='stream',form='unformatted', &
       status='replace')
  ! header
  write(u) int(jx,4) , int(iy,4) , int(kz,4)
  ! constants
  write(u) egrav , rgas , rwat , cpd , cvd , rdrcv , cpovr , rovcp , p00 , &
           ep1 , ep2 , wlhv , tzero , lrate , eomeg2 , degrad , mathpi , &
           regrav , earthrad , raddeg
  ! vertical coordinate
  write(u) zita(1:kzp1) , zitah(1:kz)
  write(u) (gzita(zita(k),mo_ztop,mo_a0), k = 1 , kzp1)
  write(u) (gzita(zitah(k),mo_ztop,mo_a0), k = 1 , kz)
  ! geometry and state over the interior
  write(u) mddom%ht(1:jx,1:iy) , mddom%hx(1:jx+1,1:iy) , mddom%hy(1:jx,1:iy+1)
  write(u) mo_atm%zeta(1:jx,1:iy,1:kz) , mo_atm%fmz(1:jx,1:iy,1:kz) , &
           mo_atm%fmzf(1:jx,1:iy,1:kzp1)
  write(u) mo_atm%t(1:jx,1:iy,1:kz) , mo_atm%qx(1:jx,1:iy,1:kz,1) , &
           mo_atm%pai(1:jx,1:iy,1:kz) , sfs%ps(1:jx,1:iy)
  close(u)
end program ref_driver
