! mod_moloch.F90 -- drop-in replacement of the reference's MOLOCH module,
! running the timestep on an MI355X through the C ABI of include/moloch_hip.h.
!
! It keeps the reference module's name and public interface
! (nompi/mod_moloch.F90:93): allocate_moloch, init_moloch and
! moloch(istep, lprint), and works on the same module state of the
! reference's own modules: mo_atm, mddom and sfs (mod_atmosphere.F90:32-34),
! dt, dx, mo_* and zita/zitah/ffilt (mod_runparams.F90:54-65), jx/iy/kz and
! the index bounds (mod_dynparam.F90).  Link it in place of
! nompi/mod_moloch.F90 together with libmoloch_hip.so.
!
! Data movement: init_moloch copies the geometry and the initial state to the
! device once.  After every moloch call the state arrays of mo_atm and sfs
! are copied back, so the module state seen by the caller is the one the
! reference would leave (copy-in/copy-out).  A caller that only reads the
! state at output time can set moloch_resident = .true. before init_moloch
! and call moloch_download() when it needs the state: the fields then stay in
! HBM between steps.
module mod_moloch

  use , intrinsic :: iso_c_binding
  use , intrinsic :: iso_fortran_env
  use mod_intkinds
  use mod_realkinds
  use mod_atmosphere
  use mod_constants
  use mod_dynparam
  use mod_runparams

  implicit none

  private

  public :: allocate_moloch , init_moloch , moloch
  public :: moloch_download , moloch_finalize
  ! ps.nc / t.nc written from the device, in place of the driver's
  ! grid_nc_create / grid_nc_write / grid_nc_destroy of sfs%ps and mo_atm%t
  public :: moloch_output_open , moloch_output_write , moloch_output_close
  ! run options, set before init_moloch
  public :: moloch_precision , moloch_halo_mpi , moloch_resident , moloch_device

  integer(c_int) , parameter :: MLC_FP64 = 0 , MLC_FP32 = 1

  integer(c_int) :: moloch_precision = MLC_FP64
  integer(c_int) :: moloch_halo_mpi = 0
  integer(c_int) :: moloch_device = 0
  logical :: moloch_resident = .false.

  ! field ids of enum mlc_field
  integer(c_int) , parameter :: F_U = 0 , F_V = 1 , F_W = 2 , F_PAI = 3 , &
    F_T = 4 , F_TETAV = 5 , F_QV = 6 , F_TVIRT = 11 , F_P = 12 , F_RHO = 13 , &
    F_QSAT = 14 , F_UX = 15 , F_VX = 16 , F_PS = 17 , F_FMZ = 18 , F_FMZF = 19 , &
    F_ZETA = 20 , F_HX = 21 , F_HY = 22 , F_MSFU = 23 , F_MSFV = 24 , &
    F_MSFX = 25 , F_ULAT = 26 , F_VLAT = 27
  integer(c_int) , parameter :: L_ZITA = 0 , L_ZITAH = 1 , L_FFILT = 2

  ! the reference's compile-time switches (nompi/mod_moloch.F90:100-104)
  integer(c_int) , parameter :: c_fulleq = 1 , c_divdamp = 1 , c_vadvtwice = 1
  integer(c_int) , parameter :: c_filterpai = 0 , c_filtertheta = 0

  type , bind(c) :: mlc_dims
    integer(c_int) :: jx , iy , kz , nqx
  end type mlc_dims

  type , bind(c) :: mlc_params
    real(c_double) :: dt , dx , dzita , ztop , zh , a0 , anu2
    integer(c_int) :: nadv , nsound , iqfrst
    integer(c_int) :: rotllr , divfilter , halo_mpi
    integer(c_int) :: fulleq , divdamp , vadvtwice , filterpai , filtertheta
    integer(c_int) :: precision , device , unfused
  end type mlc_params

  type , bind(c) :: mlc_diag
    real(c_double) :: ps_max , ps_min , u_max , u_min , v_max , v_min
    integer(c_int) :: cfl_violation
  end type mlc_diag

  interface
    integer(c_int) function mlc_default_params(p) bind(c,name='mlc_default_params')
      import :: c_int , mlc_params
      type(mlc_params) , intent(out) :: p
    end function mlc_default_params
    integer(c_int) function mlc_create(d,p,h) bind(c,name='mlc_create')
      import :: c_int , c_ptr , mlc_dims , mlc_params
      type(mlc_dims) , intent(in) :: d
      type(mlc_params) , intent(in) :: p
      type(c_ptr) , intent(out) :: h
    end function mlc_create
    integer(c_int) function mlc_set_field(h,f,a,lb,ub) bind(c,name='mlc_set_field')
      import :: c_int , c_ptr
      type(c_ptr) , value :: h
      integer(c_int) , value :: f
      type(c_ptr) , value :: a
      integer(c_int) , dimension(*) , intent(in) :: lb , ub
    end function mlc_set_field
    integer(c_int) function mlc_get_field(h,f,a,lb,ub) bind(c,name='mlc_get_field')
      import :: c_int , c_ptr
      type(c_ptr) , value :: h
      integer(c_int) , value :: f
      type(c_ptr) , value :: a
      integer(c_int) , dimension(*) , intent(in) :: lb , ub
    end function mlc_get_field
    integer(c_int) function mlc_set_levels(h,w,v,n) bind(c,name='mlc_set_levels')
      import :: c_int , c_ptr , c_double
      type(c_ptr) , value :: h
      integer(c_int) , value :: w , n
      real(c_double) , dimension(*) , intent(in) :: v
    end function mlc_set_levels
    integer(c_int) function mlc_init(h) bind(c,name='mlc_init')
      import :: c_int , c_ptr
      type(c_ptr) , value :: h
    end function mlc_init
    integer(c_int) function mlc_step(h,istep,lprint,d) bind(c,name='mlc_step')
      import :: c_int , c_ptr , mlc_diag
      type(c_ptr) , value :: h
      integer(c_int) , value :: istep , lprint
      type(mlc_diag) , intent(out) :: d
    end function mlc_step
    integer(c_int) function mlc_destroy(h) bind(c,name='mlc_destroy')
      import :: c_int , c_ptr
      type(c_ptr) , value :: h
    end function mlc_destroy
    type(c_ptr) function mlc_last_error() bind(c,name='mlc_last_error')
      import :: c_ptr
    end function mlc_last_error
    integer(c_int) function mlc_output_open(h,dir,o) bind(c,name='mlc_output_open')
      import :: c_int , c_ptr , c_char
      type(c_ptr) , value :: h
      character(kind=c_char) , dimension(*) , intent(in) :: dir
      type(c_ptr) , intent(out) :: o
    end function mlc_output_open
    integer(c_int) function mlc_output_write(o) bind(c,name='mlc_output_write')
      import :: c_int , c_ptr
      type(c_ptr) , value :: o
    end function mlc_output_write
    integer(c_int) function mlc_output_close(o) bind(c,name='mlc_output_close')
      import :: c_int , c_ptr
      type(c_ptr) , value :: o
    end function mlc_output_close
  end interface

  type(c_ptr) :: handle = c_null_ptr
  type(c_ptr) :: output = c_null_ptr

  contains

  subroutine check(rc,what)
    integer(c_int) , intent(in) :: rc
    character(len=*) , intent(in) :: what
    character(kind=c_char) , pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    if ( rc == 0 ) return
    call c_f_pointer(mlc_last_error(),msg,[512])
    text = ' '
    do n = 1 , 512
      if ( msg(n) == c_null_char ) exit
      text(n:n) = msg(n)
    end do
    call fatal(__FILE__,__LINE__,what//': '//trim(text))
    stop 1
  end subroutine check

  ! allocate_moloch (nompi/mod_moloch.F90:183-217).  The reference allocates
  ! its work arrays here; on the device they are created by init_moloch, once
  ! mo_dzita and the geometry (set by the caller after this call,
  ! nompi/regcm_dynamical_core.F90:116-120) are known.
  subroutine allocate_moloch
    implicit none
    if ( c_associated(handle) ) then
      call check(mlc_destroy(handle),'mlc_destroy')
      handle = c_null_ptr
    end if
  end subroutine allocate_moloch

  subroutine put3(f,a)
    integer(c_int) , intent(in) :: f
    real(rkx) , pointer , dimension(:,:,:) , intent(in) :: a
    integer(c_int) :: lb(3) , ub(3)
    lb = int(lbound(a),c_int)
    ub = int(ubound(a),c_int)
    call check(mlc_set_field(handle,f,c_loc(a(lb(1),lb(2),lb(3))),lb,ub),'mlc_set_field')
  end subroutine put3

  subroutine put2(f,a)
    integer(c_int) , intent(in) :: f
    real(rkx) , pointer , dimension(:,:) , intent(in) :: a
    integer(c_int) :: lb(3) , ub(3)
    lb(1:2) = int(lbound(a),c_int)
    ub(1:2) = int(ubound(a),c_int)
    lb(3) = 1
    ub(3) = 1
    call check(mlc_set_field(handle,f,c_loc(a(lb(1),lb(2))),lb,ub),'mlc_set_field')
  end subroutine put2

  subroutine get3(f,a)
    integer(c_int) , intent(in) :: f
    real(rkx) , pointer , dimension(:,:,:) , intent(inout) :: a
    integer(c_int) :: lb(3) , ub(3)
    lb = int(lbound(a),c_int)
    ub = int(ubound(a),c_int)
    call check(mlc_get_field(handle,f,c_loc(a(lb(1),lb(2),lb(3))),lb,ub),'mlc_get_field')
  end subroutine get3

  subroutine get2(f,a)
    integer(c_int) , intent(in) :: f
    real(rkx) , pointer , dimension(:,:) , intent(inout) :: a
    integer(c_int) :: lb(3) , ub(3)
    lb(1:2) = int(lbound(a),c_int)
    ub(1:2) = int(ubound(a),c_int)
    lb(3) = 1
    ub(3) = 1
    call check(mlc_get_field(handle,f,c_loc(a(lb(1),lb(2))),lb,ub),'mlc_get_field')
  end subroutine get2

  ! species n of qx, a contiguous 3-D slice
  subroutine qx_field(n,put)
    integer , intent(in) :: n
    logical , intent(in) :: put
    integer(c_int) :: lb(3) , ub(3)
    integer(c_int) :: f
    lb = int([lbound(mo_atm%qx,1),lbound(mo_atm%qx,2),lbound(mo_atm%qx,3)],c_int)
    ub = int([ubound(mo_atm%qx,1),ubound(mo_atm%qx,2),ubound(mo_atm%qx,3)],c_int)
    f = F_QV + int(n-1,c_int)
    if ( put ) then
      call check(mlc_set_field(handle,f, &
        c_loc(mo_atm%qx(lb(1),lb(2),lb(3),n)),lb,ub),'mlc_set_field')
    else
      call check(mlc_get_field(handle,f, &
        c_loc(mo_atm%qx(lb(1),lb(2),lb(3),n)),lb,ub),'mlc_get_field')
    end if
  end subroutine qx_field

  ! init_moloch (nompi/mod_moloch.F90:219-268): create the device state,
  ! copy the geometry and the initial state, derive the time-invariant
  ! coefficients (Coriolis, map-factor ratios, gzita, xknu) on the device.
  subroutine init_moloch
    implicit none
    type(mlc_dims) :: d
    type(mlc_params) :: p
    integer :: n
    call allocate_moloch
    call check(mlc_default_params(p),'mlc_default_params')
    p%dt = dt
    p%dx = dx
    p%dzita = mo_dzita
    p%ztop = mo_ztop
    p%zh = mo_h
    p%a0 = mo_a0
    p%anu2 = mo_anu2
    p%nadv = mo_nadv
    p%nsound = mo_nsound
    p%iqfrst = iqfrst
    p%rotllr = merge(1,0,iproj == 'ROTLLR')
    p%divfilter = merge(1,0,mo_divfilter)
    p%halo_mpi = moloch_halo_mpi
    p%fulleq = c_fulleq
    p%divdamp = c_divdamp
    p%vadvtwice = c_vadvtwice
    p%filterpai = c_filterpai
    p%filtertheta = c_filtertheta
    p%precision = moloch_precision
    p%device = moloch_device
    d%jx = jx
    d%iy = iy
    d%kz = kz
    d%nqx = nqx
    call check(mlc_create(d,p,handle),'mlc_create')
    ! geometry
    call put3(F_FMZ,mo_atm%fmz)
    call put3(F_FMZF,mo_atm%fmzf)
    call put3(F_ZETA,mo_atm%zeta)
    call put2(F_HX,mddom%hx)
    call put2(F_HY,mddom%hy)
    call put2(F_MSFU,mddom%msfu)
    call put2(F_MSFV,mddom%msfv)
    call put2(F_MSFX,mddom%msfx)
    call put2(F_ULAT,mddom%ulat)
    call put2(F_VLAT,mddom%vlat)
    call check(mlc_set_levels(handle,L_ZITA,zita,int(kzp1,c_int)),'mlc_set_levels')
    call check(mlc_set_levels(handle,L_ZITAH,zitah,int(kz,c_int)),'mlc_set_levels')
    call check(mlc_set_levels(handle,L_FFILT,ffilt,int(kz,c_int)),'mlc_set_levels')
    ! state
    call put3(F_U,mo_atm%u)
    call put3(F_V,mo_atm%v)
    call put3(F_W,mo_atm%w)
    call put3(F_PAI,mo_atm%pai)
    call put3(F_T,mo_atm%t)
    call put3(F_TETAV,mo_atm%tetav)
    do n = 1 , nqx
      call qx_field(n,.true.)
    end do
    call put3(F_UX,mo_atm%ux)
    call put3(F_VX,mo_atm%vx)
    call put2(F_PS,sfs%ps)
    call check(mlc_init(handle),'mlc_init')
    ! init_moloch sets w(:,:,1) = 0 (:260-262)
    mo_atm%w(:,:,1) = d_zero
  end subroutine init_moloch

  ! Copy the device state back into mo_atm and sfs.
  subroutine moloch_download
    implicit none
    integer :: n
    call get3(F_U,mo_atm%u)
    call get3(F_V,mo_atm%v)
    call get3(F_W,mo_atm%w)
    call get3(F_PAI,mo_atm%pai)
    call get3(F_T,mo_atm%t)
    call get3(F_TETAV,mo_atm%tetav)
    do n = 1 , nqx
      call qx_field(n,.false.)
    end do
    call get3(F_TVIRT,mo_atm%tvirt)
    call get3(F_P,mo_atm%p)
    call get3(F_RHO,mo_atm%rho)
    call get3(F_QSAT,mo_atm%qs)
    call get3(F_UX,mo_atm%ux)
    call get3(F_VX,mo_atm%vx)
    call get2(F_PS,sfs%ps)
  end subroutine moloch_download

  subroutine moloch_finalize
    implicit none
    if ( c_associated(handle) ) call check(mlc_destroy(handle),'mlc_destroy')
    handle = c_null_ptr
  end subroutine moloch_finalize

  ! Create ps.nc and t.nc in directory dir, as the driver's first
  ! grid_nc_create calls do (nompi/regcm_dynamical_core.F90:205-208).
  subroutine moloch_output_open(dir)
    implicit none
    character(len=*) , intent(in) :: dir
    call check(mlc_output_open(handle,trim(dir)//c_null_char,output),'mlc_output_open')
  end subroutine moloch_output_open

  ! One record of ps and t as the device holds them now; the call returns
  ! before the record reaches the files (grid_nc_write, :209-214).
  subroutine moloch_output_write
    implicit none
    call check(mlc_output_write(output),'mlc_output_write')
  end subroutine moloch_output_write

  ! Wait for the queued records and close the files (grid_nc_destroy, :220-221).
  subroutine moloch_output_close
    implicit none
    if ( c_associated(output) ) call check(mlc_output_close(output),'mlc_output_close')
    output = c_null_ptr
  end subroutine moloch_output_close

  ! moloch(istep, lprint) (nompi/mod_moloch.F90:272-421)
  subroutine moloch(istep,lprint)
    implicit none
    integer(ik4) , intent(in) :: istep
    logical , intent(in) :: lprint
    type(mlc_diag) :: dg
    call check(mlc_step(handle,int(istep,c_int),merge(1_c_int,0_c_int,lprint),dg),'mlc_step')
    if ( .not. moloch_resident ) call moloch_download
    if ( lprint ) then
      if ( dg%cfl_violation /= 0 ) then
        write (error_unit,*) 'CFL violation detected: ps max/min is not finite'
        write (error_unit,*) 'Decrease dt.'
        call fatal(__FILE__,__LINE__,'CFL VIOLATION')
      end if
      write(output_unit,*) '$$$  STEP IS : ', istep
      write(output_unit,'(a,2f8.2)') &
          ' $$$ max, min of ps (mb) = ', dg%ps_max*d_r100 , dg%ps_min*d_r100
      write(output_unit,'(a,2f8.2)') &
          ' $$$ max, min of u (m/s) = ', dg%u_max , dg%u_min
      write(output_unit,'(a,2f8.2)') &
          ' $$$ max, min of v (m/s) = ', dg%v_max , dg%v_min
    end if
  end subroutine moloch

end module mod_moloch
