! ORACLE HARNESS -- test infrastructure, not product code.
!
! A C-callable driver around the *reference* hybgen (HYCOM-src hybgen.F90,
! dpudpv.F90 and the modules they use), compiled by oracle/ref/Makefile
! straight from the HYCOM-src sources.  It loads the reference's module
! state (mod_dimensions, mod_xc, mod_cb_arrays) from caller arrays laid out
! exactly like the hyb_params / hyb_fields structs of
! include/hycom_hybgen.h, calls  hybgen(m,n,raflag)  and copies the fields
! back.  Used only by tests/ (to pin the C restatement in
! oracle/hybgen_oracle.c) and by bench.py's cpu_baseline leg.
!
! Module state set here, with the reference line that defines it:
!   mod_dimensions.F90:45-68  itdm,jtdm,idm,jdm,kdm,kk,jblk,i0,j0,ii,jj
!   mod_dimensions.F90:73     ip,iu,iv (only masks hybgen/dpudpv read)
!   mod_xc.F90:33,40          nreg, vland
!   mod_cb_arrays.F90         fields and hybgen parameters
!   mod_pipe.F90:6            lpipe (left uninitialised by the reference)
      module hybgen_ref_driver
      use iso_c_binding
      use mod_xc
      use mod_cb_arrays
      use mod_pipe, only: lpipe
      implicit none

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

      type, bind(C) :: hyb_fields_t
        type(c_ptr) :: dp, dpo, dpu, dpv, temp, saln, th3d, u, v, tracer
        type(c_ptr) :: p, pu, pv, theta, depths, topiso, depthu, depthv
        type(c_ptr) :: ubavg, vbavg, ip, iu, iv
      end type

      integer, save :: a_idm=-1, a_jdm=-1, a_kdm=-1, a_ntr=-1

      contains

      subroutine cin(dst, src, n)
      integer :: n
      real :: dst(n)
      type(c_ptr) :: src
      real(c_double), pointer :: s(:)
      if (.not. c_associated(src)) return
      call c_f_pointer(src, s, [n])
      dst = s
      end subroutine cin

      subroutine cout(src, dst, n)
      integer :: n
      real :: src(n)
      type(c_ptr) :: dst
      real(c_double), pointer :: d(:)
      if (.not. c_associated(dst)) return
      call c_f_pointer(dst, d, [n])
      d = src
      end subroutine cout

      subroutine iin(dst, src, n)
      integer :: n
      integer :: dst(n)
      type(c_ptr) :: src
      integer(c_int32_t), pointer :: s(:)
      call c_f_pointer(src, s, [n])
      dst = s
      end subroutine iin

      subroutine ensure_alloc(prm)
      type(hyb_params_t), intent(in) :: prm
      integer :: ntr
      ntr = max(prm%ntracr, prm%itracr801, prm%itracr802, prm%itracr803, 1)
      if (a_idm.eq.prm%idm .and. a_jdm.eq.prm%jdm .and. &
          a_kdm.eq.prm%kdm .and. a_ntr.eq.ntr) return
      if (a_idm.gt.0) then
        deallocate(ip, iu, iv)
        deallocate(u, v, dp, dpo, dpu, dpv, temp, saln, th3d, tracer)
        deallocate(p, pu, pv, theta, depths, topiso, depthu, depthv)
        deallocate(ubavg, vbavg, dp0k, ds0k, dx0k)
      endif
      itdm = prm%idm
      jtdm = prm%jdm
      idm  = prm%idm
      jdm  = prm%jdm
      kdm  = prm%kdm
      kk   = prm%kdm
      jblk = (jdm+2*nbdy+mxthrd-1)/mxthrd
      allocate(ip(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               iu(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               iv(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy))
      allocate( u(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
                v(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
               dp(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
              dpo(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
              dpu(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
              dpv(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
             temp(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
             saln(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
             th3d(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2), &
           tracer(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm,2,ntr))
      allocate( p(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm+1), &
               pu(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm+1), &
               pv(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm+1), &
            theta(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,kdm), &
           depths(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
           topiso(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
           depthu(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
           depthv(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
            ubavg(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,3), &
            vbavg(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy,3))
      allocate(dp0k(kdm), ds0k(kdm), dx0k(kdm))
      u = 0.0; v = 0.0; dp = 0.0; dpo = 0.0; dpu = 0.0; dpv = 0.0
      temp = 0.0; saln = 0.0; th3d = 0.0; tracer = 0.0
      a_idm = prm%idm
      a_jdm = prm%jdm
      a_kdm = prm%kdm
      a_ntr = ntr
      end subroutine ensure_alloc

      subroutine set_params(prm)
      type(hyb_params_t), intent(in) :: prm
      integer :: k
      lp     = 6
      mnproc = 1
      mproc  = 1
      nproc  = 1
      ipr    = 1
      jpr    = 1
      ijpr   = 1
      i0     = 0
      j0     = 0
      ii     = prm%ii
      jj     = prm%jj
      nreg   = prm%nreg
      vland  = prm%vland
      lpipe  = .false.
      thbase  = prm%thbase
      qhybrlx = prm%qhybrlx
      hybiso  = prm%hybiso
      hybthn  = prm%hybthn
      hybthk  = prm%hybthk
      thkbot  = prm%thkbot
      dp00i   = prm%dp00i
      dpns    = prm%dpns
      dsns    = prm%dsns
      baclin  = prm%baclin
      ra2fac  = prm%ra2fac
      nhybrd  = prm%nhybrd
      nsigma  = prm%nsigma
      hybmap  = prm%hybmap
      hybflg  = prm%hybflg
      ntracr  = prm%ntracr
      mtracr  = max(prm%itracr801, prm%itracr802, prm%itracr803, &
                    prm%ntracr) - prm%ntracr
      trcflg(:) = 0
      do k= 1,prm%ntracr
        trcflg(k) = prm%trcflg(k)
      enddo
      itracr(:)   = 0
      itracr(801) = prm%itracr801
      itracr(802) = prm%itracr802
      itracr(803) = prm%itracr803
      isopcm = prm%isopcm .ne. 0
      mxlmy  = .false.
      do k= 1,kdm
        dp0k(k) = prm%dp0k(k)
        ds0k(k) = prm%ds0k(k)
        dx0k(k) = prm%dx0k(k)
      enddo
      end subroutine set_params

! Loads the module state from the caller's arrays (ensure_alloc and
! set_params first); store_state copies the fields hybgen writes back.
      subroutine load_state(prm, fld)
      type(hyb_params_t), intent(in) :: prm
      type(hyb_fields_t), intent(in) :: fld
      integer :: n2, n3, n4, n5, n3p, ntr
      ntr = max(prm%ntracr, prm%itracr801, prm%itracr802, prm%itracr803, 1)
      n2  = (idm+2*nbdy)*(jdm+2*nbdy)
      n3  = n2*kdm
      n3p = n2*(kdm+1)
      n4  = n3*2
      n5  = n4*ntr
      call iin(ip, fld%ip, n2)
      call iin(iu, fld%iu, n2)
      call iin(iv, fld%iv, n2)
      call cin(dp,   fld%dp,   n4)
      call cin(dpo,  fld%dpo,  n4)
      call cin(dpu,  fld%dpu,  n4)
      call cin(dpv,  fld%dpv,  n4)
      call cin(temp, fld%temp, n4)
      call cin(saln, fld%saln, n4)
      call cin(th3d, fld%th3d, n4)
      call cin(u,    fld%u,    n4)
      call cin(v,    fld%v,    n4)
      if (max(prm%ntracr, prm%itracr801, prm%itracr802, prm%itracr803) &
          .gt. 0) call cin(tracer, fld%tracer, n5)
      call cin(p,  fld%p,  n3p)
      call cin(pu, fld%pu, n3p)
      call cin(pv, fld%pv, n3p)
      call cin(theta,  fld%theta,  n3)
      call cin(depths, fld%depths, n2)
      call cin(topiso, fld%topiso, n2)
      call cin(depthu, fld%depthu, n2)
      call cin(depthv, fld%depthv, n2)
      call cin(ubavg,  fld%ubavg,  n2*3)
      call cin(vbavg,  fld%vbavg,  n2*3)
      end subroutine load_state

      subroutine store_state(prm, fld)
      type(hyb_params_t), intent(in) :: prm
      type(hyb_fields_t), intent(in) :: fld
      integer :: n2, n3, n4, n5, n3p, ntr
      ntr = max(prm%ntracr, prm%itracr801, prm%itracr802, prm%itracr803, 1)
      n2  = (idm+2*nbdy)*(jdm+2*nbdy)
      n3  = n2*kdm
      n3p = n2*(kdm+1)
      n4  = n3*2
      n5  = n4*ntr
      call cout(dp,   fld%dp,   n4)
      call cout(dpo,  fld%dpo,  n4)
      call cout(dpu,  fld%dpu,  n4)
      call cout(dpv,  fld%dpv,  n4)
      call cout(temp, fld%temp, n4)
      call cout(saln, fld%saln, n4)
      call cout(th3d, fld%th3d, n4)
      call cout(u,    fld%u,    n4)
      call cout(v,    fld%v,    n4)
      if (max(prm%ntracr, prm%itracr801, prm%itracr802, prm%itracr803) &
          .gt. 0) call cout(tracer, fld%tracer, n5)
      call cout(p,  fld%p,  n3p)
      call cout(pu, fld%pu, n3p)
      call cout(pv, fld%pv, n3p)
      end subroutine store_state

! Runs hybgen(m,n,raflag) -- the reference's, or whatever subroutine hybgen
! the library is linked with -- on the caller's arrays, in place.  secs
! returns the wall time of the hybgen call alone (no copies).
      integer(c_int) function href_hybgen(prm, fld, m, n, raflag, secs) &
          bind(C, name='href_hybgen')
      type(hyb_params_t), intent(in) :: prm
      type(hyb_fields_t), intent(in) :: fld
      integer(c_int), value :: m, n, raflag
      real(c_double), intent(out) :: secs
      integer(8) :: t0, t1, rate
      external hybgen
      href_hybgen = -1
      secs = 0.0
      if (prm%nbdy .ne. nbdy) return
      if (prm%hybflg .ne. 0 .or. prm%mxlmy .ne. 0) return
      call ensure_alloc(prm)
      call set_params(prm)
      call load_state(prm, fld)
      call system_clock(t0, rate)
      call hybgen(int(m), int(n), raflag .ne. 0)
      call system_clock(t1)
      secs = real(t1-t0, 8)/real(rate, 8)
      call store_state(prm, fld)
      href_hybgen = 0
      end function href_hybgen

      end module hybgen_ref_driver
