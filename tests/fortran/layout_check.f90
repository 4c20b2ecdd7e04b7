! Prints fields of salsa_bin_desc / salsa_params filled by the C library
! through the BIND(C) twins of mo_salsa_amd_cabi; the test compares them with
! the values the Python ctypes binding reads from the same calls.
PROGRAM layout_check
  USE, INTRINSIC :: ISO_C_BINDING
  USE mo_salsa_amd_cabi
  IMPLICIT NONE
  TYPE(salsa_bin_desc) :: d
  TYPE(salsa_params) :: p
  INTEGER(C_INT32_T) :: ids(7)
  INTEGER(C_INT) :: rc
  ids = 0
  ids(2) = 3
  rc = salsa_bin_desc_init(d, [3_C_INT32_T, 4_C_INT32_T], [3.e-9_C_DOUBLE, 9.703e-8_C_DOUBLE, 1.e-5_C_DOUBLE, &
                           1.e-5_C_DOUBLE], 7_C_INT32_T, 20.e-6_C_DOUBLE, 4._C_DOUBLE, 7_C_INT32_T, 2.e-6_C_DOUBLE, &
                           20._C_DOUBLE, 2_C_INT32_T, ids)
  WRITE(*,'(A,I0)') 'rc ', rc
  WRITE(*,'(A,7(1X,I0))') 'idx ', d%in1a, d%fn1a, d%in2a, d%fn2a, d%in2b, d%fn2b, d%nbins
  WRITE(*,'(A,4(1X,I0))') 'par ', d%ica%cur, d%ica%par, d%fcb%cur, d%fcb%par
  WRITE(*,'(A,5(1X,I0))') 'cnt ', d%ncld, d%ntotal, d%nspec, d%iwa, d%irim
  WRITE(*,'(A,3(1X,ES24.16))') 'dbl ', d%dmid(5), d%precpbins(3), d%rholiq(2)
  WRITE(*,'(A,2(1X,I0))') 'ph ', d%phase(12), d%phase(d%ntotal)
  rc = salsa_params_default(p, 5_C_INT32_T)
  WRITE(*,'(A,I0)') 'prc ', rc
  WRITE(*,'(A,6(1X,I0))') 'psw ', p%level, p%lscgii, p%lscndh2oic, p%lsfreeRH, p%precision, p%hxy
  WRITE(*,'(A,4(1X,ES24.16))') 'pdb ', p%rhlim, p%ice_sigma, p%charge_coll_enh, p%cgintvl
  p%lsdistupdate = 7
  p%lsactiv = 9
  WRITE(*,'(A,3(1X,I0))') 'pmd ', p%lscoag_mode, p%lsdistupdate, p%lsactiv
END PROGRAM layout_check
