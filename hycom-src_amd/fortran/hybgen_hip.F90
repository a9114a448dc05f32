! hybgen_hip.F90 -- drop-in replacement of HYCOM's  subroutine hybgen
! (hybgen.F90:14-306): same name, same arguments, same module state.  Link
! this file instead of hybgen.F90 and link libhycom_hybgen.so; the callers
! (mod_hycom.F90:2650, inicon.F90:544,559) stay unchanged.
!
! The call hands HYCOM's own module arrays (mod_cb_arrays, mod_dimensions)
! to the C ABI of include/hycom_hybgen.h as plain pointers, in their Fortran
! layout, plus the hybgen parameters as one struct; the library computes
! hybgen on the GPU and writes the results back into the same arrays.
! Errors stop the model through xcstop, as the reference's own checks do
! (blkdat.F90:946-1012, dpudpv.F90:47-55).
!
! Built with the model's flags (REAL8: real is 8 bytes; default integer is
! 4 bytes).  HYBGEN_DENFIX selects the same variant as in hybgen.F90:319-325.
      module hybgen_hip_abi
      use iso_c_binding
      implicit none
      private
      public :: hyb_params_t, hyb_fields_t, hyb_pack, &
                hycom_hybgen_c, hycom_hybgen_last_error_c

! include/hycom_hybgen.h: hyb_params
      type, bind(C) :: hyb_params_t
        integer(c_int32_t) :: idm, jdm, ii, jj, nbdy, kdm, ntracr, &
                              nhybrd, nsigma, hybmap, hybflg, nreg, &
                              isopcm, mxlmy, ldenfix, &
                              itracr801, itracr802, itracr803, &
                              precision, check_finite
        integer(c_int32_t) :: trcflg(99)
        real(c_double)     :: thbase, qhybrlx, hybiso, hybthn, hybthk, &
                              thkbot, dp00i, dpns, dsns, baclin, ra2fac, &
                              vland
        real(c_double)     :: dp0k(64), ds0k(64), dx0k(64)
      end type

! include/hycom_hybgen.h: hyb_fields
      type, bind(C) :: hyb_fields_t
        type(c_ptr) :: dp, dpo, dpu, dpv, temp, saln, th3d, u, v, tracer
        type(c_ptr) :: p, pu, pv, theta, depths, topiso, depthu, depthv
        type(c_ptr) :: ubavg, vbavg, ip, iu, iv
      end type

      interface
        integer(c_int) function hycom_hybgen_c(prm, fld, m, n, raflag) &
            bind(C, name='hycom_hybgen')
          import :: c_int, hyb_params_t, hyb_fields_t
          type(hyb_params_t), intent(in) :: prm
          type(hyb_fields_t), intent(in) :: fld
          integer(c_int), value :: m, n, raflag
        end function
        type(c_ptr) function hycom_hybgen_last_error_c() &
            bind(C, name='hycom_hybgen_last_error')
          import :: c_ptr
        end function
      end interface

      contains

! address of the first element of a contiguous module array
      type(c_ptr) function raddr(x)
      real, target, intent(in) :: x(*)
      raddr = c_loc(x(1))
      end function raddr

      type(c_ptr) function iaddr(x)
      integer, target, intent(in) :: x(*)
      iaddr = c_loc(x(1))
      end function iaddr

! the hybgen parameters and fields of the model, for the C ABI
      subroutine hyb_pack(prm, fld)
      use mod_xc          ! idm, jdm, ii, jj, nbdy, kdm, nreg, vland
      use mod_cb_arrays   ! fields and hybgen parameters
      type(hyb_params_t), intent(out) :: prm
      type(hyb_fields_t), intent(out) :: fld
      integer :: k
      prm%idm    = idm
      prm%jdm    = jdm
      prm%ii     = ii
      prm%jj     = jj
      prm%nbdy   = nbdy
      prm%kdm    = kdm
      prm%ntracr = ntracr
      prm%nhybrd = nhybrd
      prm%nsigma = nsigma
      prm%hybmap = hybmap
      prm%hybflg = hybflg
      prm%nreg   = nreg
      prm%isopcm = merge(1, 0, isopcm)
      prm%mxlmy  = merge(1, 0, mxlmy)
#if defined(HYBGEN_DENFIX)
      prm%ldenfix = 1
#else
      prm%ldenfix = 0
#endif
      prm%itracr801 = itracr(801)
      prm%itracr802 = itracr(802)
      prm%itracr803 = itracr(803)
! HYCOM_HYBGEN_PRECISION=mixed in the environment selects the mixed form
! (HYB_PREC_MIXED); otherwise FP64, bitwise the reference's
      prm%precision    = hybgen_precision()
      prm%check_finite = 0
      prm%trcflg(:) = 0
      do k= 1,min(ntracr,99)
        prm%trcflg(k) = trcflg(k)
      enddo
      prm%thbase  = thbase
      prm%qhybrlx = qhybrlx
      prm%hybiso  = hybiso
      prm%hybthn  = hybthn
      prm%hybthk  = hybthk
      prm%thkbot  = thkbot
      prm%dp00i   = dp00i
      prm%dpns    = dpns
      prm%dsns    = dsns
      prm%baclin  = baclin
      prm%ra2fac  = ra2fac
      prm%vland   = vland
      prm%dp0k(:) = 0.0
      prm%ds0k(:) = 0.0
      prm%dx0k(:) = 0.0
      do k= 1,min(kdm,64)
        prm%dp0k(k) = dp0k(k)
        prm%ds0k(k) = ds0k(k)
        prm%dx0k(k) = dx0k(k)
      enddo
      fld%dp     = raddr(dp)
      fld%dpo    = raddr(dpo)
      fld%dpu    = raddr(dpu)
      fld%dpv    = raddr(dpv)
      fld%temp   = raddr(temp)
      fld%saln   = raddr(saln)
      fld%th3d   = raddr(th3d)
      fld%u      = raddr(u)
      fld%v      = raddr(v)
      if (ntracr+mtracr.gt.0) then  ! mtracr: the diagnostic tracers
        fld%tracer = raddr(tracer)
      else
        fld%tracer = c_null_ptr
      endif
      fld%p      = raddr(p)
      fld%pu     = raddr(pu)
      fld%pv     = raddr(pv)
      fld%theta  = raddr(theta)
      fld%depths = raddr(depths)
      fld%topiso = raddr(topiso)
      fld%depthu = raddr(depthu)
      fld%depthv = raddr(depthv)
      fld%ubavg  = raddr(ubavg)
      fld%vbavg  = raddr(vbavg)
      fld%ip     = iaddr(ip)
      fld%iu     = iaddr(iu)
      fld%iv     = iaddr(iv)
      end subroutine hyb_pack

! 1 when HYCOM_HYBGEN_PRECISION is 'mixed', else 0; read on the first call
      integer function hybgen_precision()
      integer, save :: prec = -1
      character(len=16) :: val
      integer :: lval, ios
      if (prec.lt.0) then
        prec = 0
        call get_environment_variable('HYCOM_HYBGEN_PRECISION', val, lval, ios)
        if (ios.eq.0 .and. lval.gt.0) then
          if (val(1:lval).eq.'mixed') prec = 1
        endif
      endif
      hybgen_precision = prec
      end function hybgen_precision

      end module hybgen_hip_abi

! hybgen.F90:14-23: the entry the model calls
      subroutine hybgen(m,n, hybgen_raflag)
      use iso_c_binding
      use mod_xc
      use hybgen_hip_abi
      implicit none
      integer, intent(in) :: m,n
      logical, intent(in) :: hybgen_raflag
      type(hyb_params_t) :: prm
      type(hyb_fields_t) :: fld
      integer(c_int) :: rc
      character(kind=c_char), pointer :: msg(:)
      integer :: k
      call hyb_pack(prm, fld)
      rc = hycom_hybgen_c(prm, fld, m, n, merge(1, 0, hybgen_raflag))
      if (rc.ne.0) then
        if (mnproc.eq.1) then
          call c_f_pointer(hycom_hybgen_last_error_c(), msg, [256])
          write(lp,'(a,i3,1x,256a)') 'error - hybgen_hip rc=', rc, &
            (msg(k), k= 1,index_nul(msg))
          call flush(lp)
        endif
        call xcstop('(hybgen_hip)')
        stop '(hybgen_hip)'
      endif
      contains
      integer function index_nul(s)
      character(kind=c_char) :: s(:)
      integer :: i
      index_nul = size(s)
      do i= 1,size(s)
        if (s(i).eq.c_null_char) then
          index_nul = i-1
          return
        endif
      enddo
      end function index_nul
      end subroutine hybgen
