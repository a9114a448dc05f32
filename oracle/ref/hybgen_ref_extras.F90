! ORACLE HARNESS -- test infrastructure, not product code.
!
! C-callable pieces of the reference besides hybgen itself: the equation
! of state, the constants and one column through a remapper pair.  Kept
! apart from hybgen_ref_driver.F90 so that the shim check library, which
! links the shim's hybgen instead of hybgen.F90, does not need them.
      module hybgen_ref_extras
      use iso_c_binding
      use mod_xc
      use mod_cb_arrays
      implicit none
      contains

! sig(t,s) of the reference equation of state (stmt_fns.h:505-511).
      real(c_double) function href_sig(xt, xs) bind(C, name='href_sig')
      real(c_double), value :: xt, xs
#include "stmt_fns.h"
      href_sig = sig(xt, xs)
      end function href_sig

! The reference's physical constants (mod_cb_arrays.F90:843-857) and the
! folded EOS coefficients (stmt_fns.h:251-262), for bitwise pinning.
      subroutine href_consts(c) bind(C, name='href_consts')
      real(c_double), intent(out) :: c(16)
#include "stmt_fns.h"
      c(1) = onem
      c(2) = tenm
      c(3) = onemm
      c(4) = epsil
      c(5) = qonem
      c(6) = svref
      c(7) = c101
      c(8) = c103
      c(9) = c105
      c(10) = c108
      c(11) = c109
      c(12) = c111
      c(13) = rpdb
      c(14) = 0.001*onemm
      c(15) = 0.0
      c(16) = 0.0
      end subroutine href_consts

! One column through the reference remapper pair, as hybgenaj calls them
! (hybgen.F90:1530-1548): method 0 PCM, 1 PLM, 2 PPM, 3 WENO.
      subroutine href_remap(method, kin, si, pin, po, lcm_i, thin, so) &
          bind(C, name='href_remap')
      integer(c_int), value :: method, kin
      real(c_double), intent(in) :: si(kin), pin(kin+1), po(kin+1)
      integer(c_int), intent(in) :: lcm_i(kin)
      real(c_double), value :: thin
      real(c_double), intent(out) :: so(kin)
      real :: dprs(kin), dpi(kin), c1d(kin,1,3), s1(kin,1), f1(kin,1)
      real :: pi1(kin+1), po1(kin+1), th
      logical :: lcm(kin)
      integer :: k
      external hybgen_pcm_remap, hybgen_plm_coefs, hybgen_plm_remap, &
               hybgen_ppm_coefs, hybgen_ppm_remap, &
               hybgen_weno_coefs, hybgen_weno_remap
      th = thin
      do k= 1,kin
        s1(k,1) = si(k)
        lcm(k) = lcm_i(k) .ne. 0
      enddo
      do k= 1,kin+1
        pi1(k) = pin(k)
        po1(k) = po(k)
      enddo
      do k= 1,kin
        dprs(k) = pi1(k+1)-pi1(k)
        dpi(k)  = max(dprs(k),th)
      enddo
      if     (method.eq.0) then
        call hybgen_pcm_remap(s1,pi1,dprs,f1,po1,kin,kin,1,th)
      elseif (method.eq.1) then
        call hybgen_plm_coefs(s1,dprs,lcm,c1d,kin,1,th)
        call hybgen_plm_remap(s1,pi1,dprs,c1d,f1,po1,kin,kin,1,th)
      elseif (method.eq.2) then
        call hybgen_ppm_coefs(s1,dpi,lcm,c1d,kin,1,th)
        call hybgen_ppm_remap(s1,pi1,dprs,c1d,f1,po1,kin,kin,1,th)
      else
        call hybgen_weno_coefs(s1,dpi,lcm,c1d,kin,1,th)
        call hybgen_weno_remap(s1,pi1,dprs,c1d,f1,po1,kin,kin,1,th)
      endif
      do k= 1,kin
        so(k) = f1(k,1)
      enddo
      end subroutine href_remap

! One column of the remap stokes_vertical_j makes (mod_stokes.F90:686-694):
! ks fields of ki fixed layers onto ko model layers, hybgen_plm_coefs on
! dpi = max(dprs,thin) with no PCM layers, then hybgen_plm_remap.  The
! blend and the interfaces are the caller's (si, pin, po).
      subroutine href_plm_dpi_column(ki, ko, ks, si, pin, po, thin, so) &
          bind(C, name='href_plm_dpi_column')
      integer(c_int), value :: ki, ko, ks
      real(c_double), intent(in) :: si(ki,ks), pin(ki+1), po(ko+1)
      real(c_double), value :: thin
      real(c_double), intent(out) :: so(ko,ks)
      real :: s1(ki,ks), f1(ko,ks), c1d(ki,ks), dprs(ki), dpi(ki)
      real :: pi1(ki+1), po1(ko+1), th
      logical :: lcm(ki)
      integer :: k
      external hybgen_plm_coefs, hybgen_plm_remap
      th = thin
      s1 = si
      pi1 = pin
      po1 = po
      lcm(:) = .false.
      do k= 1,ki
        dprs(k) = pi1(k+1)-pi1(k)
        dpi(k)  = max(dprs(k),th)
      enddo
      call hybgen_plm_coefs(s1,dpi,lcm,c1d,ki,ks,th)
      call hybgen_plm_remap(s1,pi1,dprs,c1d,f1,po1,ki,ko,ks,th)
      so = f1
      end subroutine href_plm_dpi_column

! .a records through the reference's own array I/O (mod_za.F90 with
! mod_za_sm.h): mode 1 writes h with zaiowr3 (real*4, 2^100 padding, byte
! order swapped by zaio_endian under ENDIAN_IO), mode 0 reads it back with
! zaiord3; both return each record's range.  zaiost allocates its buffer
! for the extents of the first call, so every call must use those extents
! (-2 otherwise).
      integer(c_int) function href_za(mode, path, plen, idm_, jdm_, nrec, h, hmin, hmax) &
          bind(C, name='href_za')
      use mod_za
      use hybgen_ref_driver, only: hyb_params_t, ensure_alloc
      integer(c_int), value :: mode, plen, idm_, jdm_, nrec
      character(kind=c_char), intent(in) :: path(plen)
      real(c_double), intent(inout) :: h(idm_+2*nbdy, jdm_+2*nbdy, nrec)
      real(c_double), intent(out) :: hmin(nrec), hmax(nrec)
      type(hyb_params_t) :: prm
      character(len=512) :: cfile
      integer, allocatable :: msk(:,:)
      real :: h8(idm_+2*nbdy, jdm_+2*nbdy, nrec), xmin(nrec), xmax(nrec)
      integer :: k
      logical, save :: ready = .false.
      integer, save :: rid = -1, rjd = -1
      href_za = -1
      if (plen .gt. 512) return
      cfile = ' '
      do k= 1,plen
        cfile(k:k) = path(k)
      enddo
      prm%idm = idm_
      prm%jdm = jdm_
      prm%kdm = 1
      prm%ntracr = 0
      prm%itracr801 = 0
      prm%itracr802 = 0
      prm%itracr803 = 0
      call ensure_alloc(prm)
      lp = 6
      mnproc = 1
      if (.not. ready) then
        call zaiost
        ready = .true.
        rid = idm_
        rjd = jdm_
      elseif (rid .ne. idm_ .or. rjd .ne. jdm_) then
        href_za = -2
        return
      endif
      allocate(msk(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy))
      msk = 1
      h8 = h
      if (mode .eq. 1) then
        call zaiopf(cfile(1:plen), 'new', 77)
        call zaiowr3(h8, nrec, msk, .false., xmin, xmax, 77, .true.)
      else
        call zaiopf(cfile(1:plen), 'old', 77)
        call zaiord3(h8, nrec, msk, .false., xmin, xmax, 77)
      endif
      call zaiocl(77)
      h = h8
      hmin = xmin
      hmax = xmax
      href_za = 0
      end function href_za

! Every array gindex_allocate (mod_dimensions.F90:118-210) allocates,
! released first so that it can allocate them again at the current
! extents (ensure_alloc allocates ip, iu, iv on its own).
      subroutine gindex_reset
      if (allocated(ip)) deallocate(ip)
      if (allocated(iu)) deallocate(iu)
      if (allocated(iv)) deallocate(iv)
      if (allocated(iq)) deallocate(iq, iuopn, ivopn, ishlf, ipa, &
          ipim1x, ipip1x, ipjm1x, ipjp1x, ipim1, ipip1, ipjm1, ipjp1, &
          ifp, ilp, ifq, ilq, ifu, ilu, ifv, ilv, &
          jfp, jlp, jfq, jlq, jfu, jlu, jfv, jlv, &
          isp, isq, isu, isv, jsp, jsq, jsu, jsv, &
          allip, alliq, alliu, alliv)
      call gindex_allocate
      end subroutine gindex_reset

! The model start on one tile through the reference's own routines: the
! bathymetry as geopar.F90:244-256 treats it (voids to land, then
! bigrid), geopar's zero start of depthu, depthv, dpu, dpv and p(:,:,1)
! (760-800), then mod_hycom.F90:1684-1722: dpthuv, the dp halo, and p and
! dpudpv of both time levels.  All arrays are (NI,NJ[,...]) with the
! driver's nbdy; depths, pbot and dp come back with their halos filled.
      integer(c_int) function href_start(prm, depths_, pbot_, dp_, ip_, iu_, iv_, nreg_, &
          depthu_, depthv_, p_, dpu_, dpv_) bind(C, name='href_start')
      use hybgen_ref_driver, only: hyb_params_t, ensure_alloc, set_params, cin, cout
      type(hyb_params_t), intent(in) :: prm
      type(c_ptr), value :: depths_, pbot_, dp_, ip_, iu_, iv_
      type(c_ptr), value :: depthu_, depthv_, p_, dpu_, dpv_
      integer(c_int32_t), intent(out) :: nreg_
      real, allocatable :: util1(:,:), util2(:,:), util3(:,:)
      integer(c_int32_t), pointer :: io(:)
      integer :: i, j, k, nm, n2, margin
      href_start = -1
      if (prm%nbdy .ne. nbdy) return
      call ensure_alloc(prm)
      call set_params(prm)
      call gindex_reset
      if (allocated(pbot)) deallocate(pbot, pbotmin, corio, pvtrop)
      allocate(pbot(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               pbotmin(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               corio(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               pvtrop(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy))
      allocate(util1(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               util2(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy), &
               util3(1-nbdy:idm+nbdy,1-nbdy:jdm+nbdy))
      n2 = (idm+2*nbdy)*(jdm+2*nbdy)
      util1 = 0.0
      util2 = 0.0
      util3 = 0.0
      call cin(depths, depths_, n2)
      do j= 1,jj
        do i= 1,ii
          if (depths(i,j).gt.0.5*hugel) depths(i,j) = 0.0
        enddo
      enddo
      nreg  = -1   ! TYPE=one: bigrid decides
      vland = 0.0  ! xcspmd's value until geopar sets it after bigrid
      call bigrid(depths, 0, util1, util2, util3)
      nreg_ = nreg
      vland = prm%vland
      depthu = 0.0
      depthv = 0.0
      dpu = 0.0
      dpv = 0.0
      p(:,:,1) = 0.0
      corio = 0.0
      oneta0 = 0.0
      lbflag = 0
      call cin(pbot, pbot_, n2)
      call dpthuv
      call cin(dp, dp_, n2*kdm*2)
      call xctilr(dp(1-nbdy,1-nbdy,1,1), 1, 2*kk, nbdy, nbdy, halo_ps)
      margin = nbdy
      do nm= 1,2
        do j= 1-margin,jj+margin
          do i= 1-margin,ii+margin
            if (ip(i,j).ne.0) then
              p(i,j,1) = 0.0
              do k= 1,kk
                p(i,j,k+1) = p(i,j,k)+dp(i,j,k,nm)
              enddo
            endif
          enddo
        enddo
        call dpudpv(dpu(1-nbdy,1-nbdy,1,nm), dpv(1-nbdy,1-nbdy,1,nm), &
                    p, depthu, depthv, margin, max(0,margin-1))
      enddo
      call cout(depths, depths_, n2)
      call cout(pbot, pbot_, n2)
      call cout(dp, dp_, n2*kdm*2)
      call cout(depthu, depthu_, n2)
      call cout(depthv, depthv_, n2)
      call cout(p, p_, n2*(kdm+1))
      call cout(dpu, dpu_, n2*kdm*2)
      call cout(dpv, dpv_, n2*kdm*2)
      call c_f_pointer(ip_, io, [n2])
      io = reshape(ip, [n2])
      call c_f_pointer(iu_, io, [n2])
      io = reshape(iu, [n2])
      call c_f_pointer(iv_, io, [n2])
      io = reshape(iv, [n2])
      href_start = 0
      end function href_start

      end module hybgen_ref_extras
