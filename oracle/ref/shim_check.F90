! ORACLE HARNESS -- test infrastructure, not product code.
!
! Checks the drop-in shim hycom-src_amd/fortran/hybgen_hip.F90: loads the
! model's module state from caller arrays (the harness of
! hybgen_ref_driver.F90) and returns what the shim's hyb_pack hands to the
! C ABI: the parameter struct and pointers to the module arrays.  The test
! then runs the oracle on those pointers and compares.
      module shim_check
      use iso_c_binding
      use hybgen_ref_driver, only: drv_params => hyb_params_t, &
                                   drv_fields => hyb_fields_t, &
                                   ensure_alloc, set_params, &
                                   load_state, store_state
      use hybgen_hip_abi, only: hyb_params_t, hyb_fields_t, hyb_pack
      implicit none
      contains

      integer(c_int) function shimchk_pack(prm_in, fld_in, prm_out, fld_out) &
          bind(C, name='shimchk_pack')
      type(drv_params), intent(in) :: prm_in
      type(drv_fields), intent(in) :: fld_in
      type(hyb_params_t), intent(out) :: prm_out
      type(hyb_fields_t), intent(out) :: fld_out
      call ensure_alloc(prm_in)
      call set_params(prm_in)
      call load_state(prm_in, fld_in)
      call hyb_pack(prm_out, fld_out)
      shimchk_pack = 0
      end function shimchk_pack

      integer(c_int) function shimchk_store(prm_in, fld_in) &
          bind(C, name='shimchk_store')
      type(drv_params), intent(in) :: prm_in
      type(drv_fields), intent(in) :: fld_in
      call store_state(prm_in, fld_in)
      shimchk_store = 0
      end function shimchk_store

      end module shim_check
