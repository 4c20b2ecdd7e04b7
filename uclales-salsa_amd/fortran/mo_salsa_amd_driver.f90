!> Drop-in run_SALSA that runs the coagulation + condensation step of SALSA
!> on the GPU through libsalsa_amd.
!>
!> The dummy argument list is that of run_SALSA in mo_salsa_driver.f90:112-133,
!> so the two call sites (step.f90:319-328 and init.f90:101-107) link unchanged
!> once their USE statement names this module instead of mo_salsa_driver.
!> The fields are fetched from the LES FieldArrays by the same names as
!> initialize_salsa_driver (mo_salsa_driver.f90:481-561) and passed as host
!> pointers; the library adds the tendencies (Prog tlev 2) and overwrites the
!> process-rate Diag fields of the interior cells.
!>
!> Only the processes of this step are provided: coagulation, condensation,
!> autoconv2 (lsauto mode 2), cloud activation (lsactiv mode 1, interstitial;
!> mode 2, cloud base, from wp into pa_nactd / pa_vactd through device copies
!> of the three fields) and distr_update.  A configuration that also switches
!> on secondary ice, ice nucleation, melting, gas-phase condensation or
!> nucleation stops with a message naming the process (mo_salsa.f90:92-153
!> calls them in the same step).
!>
!> The precision mode is taken from the environment variable
!> SALSA_AMD_PRECISION ("fp64", the default, or "mixed").
MODULE mo_salsa_amd_driver
  USE, INTRINSIC :: ISO_C_BINDING
  USE mo_salsa_amd_cabi
  USE classFieldArray, ONLY : FieldArray
  USE mo_structured_datatypes, ONLY : FloatArray3d, FloatArray4d
  USE mo_submctl, ONLY : nbin, reglim, bloPrc, bloIce, ncld, spec, rhlim, eddy_dis, Eiagg_max, Eiagg_min,  &
                         lscoag, lscnd, lsauto, lsactiv, lsicenucl, lsicemelt, lssecice, lsfreeRH,         &
                         lscgaa, lscgcc, lscgpp, lscgii, lscgca, lscgpa, lscgia, lscgpc, lscgic, lscgip,   &
                         lscndh2oae, lscndh2ocl, lscndh2oic, lscndgas, nsnucl, lsdistupdate, cgintvl
  USE mo_ice_shape, ONLY : iceShapeAlpha, iceShapeBeta, iceShapeGamma, iceShapeSigma
  USE mo_salsa_types, ONLY : allSALSA
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: run_SALSA, salsa_amd_finalize

  TYPE(C_PTR), SAVE :: ctx = C_NULL_PTR
  REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: rates(:,:,:,:)
  ! cloud-base activation: device copies of wp, pa_nactd and pa_vactd
  TYPE(C_PTR), SAVE :: d_w = C_NULL_PTR, d_nactd = C_NULL_PTR, d_vactd = C_NULL_PTR
  INTEGER(C_SIZE_T), SAVE :: act_n3 = 0

  ! the HIP runtime calls the activation fields need (hip_runtime_api.h)
  INTEGER(C_INT), PARAMETER :: hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2
  INTERFACE
     INTEGER(C_INT) FUNCTION hipMalloc(ptr, bytes) BIND(C, name="hipMalloc")
       IMPORT :: C_INT, C_PTR, C_SIZE_T
       TYPE(C_PTR)              :: ptr
       INTEGER(C_SIZE_T), VALUE :: bytes
     END FUNCTION hipMalloc
     INTEGER(C_INT) FUNCTION hipFree(ptr) BIND(C, name="hipFree")
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ptr
     END FUNCTION hipFree
     INTEGER(C_INT) FUNCTION hipMemset(ptr, val, bytes) BIND(C, name="hipMemset")
       IMPORT :: C_INT, C_PTR, C_SIZE_T
       TYPE(C_PTR), VALUE       :: ptr
       INTEGER(C_INT), VALUE    :: val
       INTEGER(C_SIZE_T), VALUE :: bytes
     END FUNCTION hipMemset
     INTEGER(C_INT) FUNCTION hipMemcpy(dst, src, bytes, kind) BIND(C, name="hipMemcpy")
       IMPORT :: C_INT, C_PTR, C_SIZE_T
       TYPE(C_PTR), VALUE       :: dst, src
       INTEGER(C_SIZE_T), VALUE :: bytes
       INTEGER(C_INT), VALUE    :: kind
     END FUNCTION hipMemcpy
  END INTERFACE

  CHARACTER(len=4), PARAMETER :: dry_names(7) = ['SO4 ', 'OC  ', 'BC  ', 'DU  ', 'SS  ', 'NO  ', 'NH  ']
  CHARACTER(len=16), PARAMETER :: rate_names(SALSA_NRATE_L5) = [CHARACTER(len=16) ::                      &
       's_m_autoc', 's_m_autoc50', 's_m_autoc80', 's_m_accr', 's_m_accr50', 's_m_accr80',                  &
       's_m_ACcoll_dry', 's_m_APcoll_dry', 's_n_activ', 's_m_conda', 's_m_condc', 's_m_condp',              &
       's_m_AIcoll_dry', 's_n_icehom', 's_n_icedep', 's_n_iceimm', 's_m_condi']

CONTAINS

  SUBROUTINE run_SALSA(Diag, Prog, nzp, nxp, nyp, ns, wp,     &
                       pa_nactd, pa_vactd, tstep, time, istp,  &
                       level, lcharge, initialize)
    TYPE(FieldArray), INTENT(inout) :: Diag, Prog
    INTEGER, INTENT(in) :: nzp, nxp, nyp, ns
    REAL, INTENT(in)    :: tstep
    REAL, INTENT(in)    :: time
    INTEGER, INTENT(in) :: istp
    LOGICAL, INTENT(in) :: initialize
    REAL, INTENT(in), TARGET  :: wp(nzp,nxp,nyp)
    INTEGER, INTENT(in) :: level
    LOGICAL, INTENT(in) :: lcharge
    REAL, INTENT(out), TARGET :: pa_vactd(nzp,nxp,nyp,ns*ncld)
    REAL, INTENT(out), TARGET :: pa_nactd(nzp,nxp,nyp,ncld)

    TYPE(FloatArray3d), POINTER :: ppress => NULL(), ptk => NULL(), prs => NULL(), prsi => NULL(), &
                                   pdn => NULL(), prv => NULL(), prt => NULL(), fld => NULL()
    TYPE(FloatArray4d), POINTER :: pnap => NULL(), pnat => NULL(), pmap => NULL(), pmat => NULL(), &
                                   pncp => NULL(), pnct => NULL(), pmcp => NULL(), pmct => NULL(), &
                                   pnpp => NULL(), pnpt => NULL(), pmpp => NULL(), pmpt => NULL(), &
                                   pnip => NULL(), pnit => NULL(), pmip => NULL(), pmit => NULL()
    TYPE(C_PTR) :: n_p(4), n_t(4), m_p(4), m_t(4), crsi
    INTEGER :: nrf, f, nwet
    INTEGER(C_INT) :: rc
    INTEGER(C_INT32_T) :: sflags
    INTEGER(C_SIZE_T) :: n3
    LOGICAL :: cloud_base

    CALL check_switches()
    IF (.NOT. C_ASSOCIATED(ctx)) CALL create_context(level, lcharge, nzp, nxp, nyp)

    ! cloud-base activation (lsactiv mode 2): wp to the device before the
    ! step, the activated droplets back after it (mo_salsa_driver.f90:360-366)
    pa_nactd = 0.
    pa_vactd = 0.
    cloud_base = lsactiv%state .AND. lsactiv%mode == 2
    nwet = spec%getNSpec(type="wet")
    n3 = INT(nzp, C_SIZE_T) * INT(nxp, C_SIZE_T) * INT(nyp, C_SIZE_T)
    IF (cloud_base) THEN
       IF (ns < nwet) CALL fail('pa_vactd holds fewer species than the wet species')
       IF (.NOT. C_ASSOCIATED(d_w)) CALL alloc_activation_fields(n3, nwet)
       rc = hipMemcpy(d_w, C_LOC(wp(1,1,1)), 8_C_SIZE_T * n3, hipMemcpyHostToDevice)
       IF (rc /= 0) CALL fail('copying wp to the device')
    END IF

    CALL Diag%getData(1, ppress, name="press")
    CALL Diag%getData(1, ptk, name="temp")
    CALL Diag%getData(1, prs, name="rsl")
    CALL Diag%getData(1, pdn, name="dn")
    CALL Prog%getData(1, prv, name="rp")
    CALL Prog%getData(2, prt, name="rp")
    CALL Prog%getData(1, pnap, name="naero");  CALL Prog%getData(2, pnat, name="naero")
    CALL Prog%getData(1, pncp, name="ncloud"); CALL Prog%getData(2, pnct, name="ncloud")
    CALL Prog%getData(1, pnpp, name="nprecp"); CALL Prog%getData(2, pnpt, name="nprecp")
    CALL Prog%getData(1, pmap, name="maero");  CALL Prog%getData(2, pmat, name="maero")
    CALL Prog%getData(1, pmcp, name="mcloud"); CALL Prog%getData(2, pmct, name="mcloud")
    CALL Prog%getData(1, pmpp, name="mprecp"); CALL Prog%getData(2, pmpt, name="mprecp")
    n_p = C_NULL_PTR; n_t = C_NULL_PTR; m_p = C_NULL_PTR; m_t = C_NULL_PTR
    n_p(1) = C_LOC(pnap%d(1,1,1,1)); n_t(1) = C_LOC(pnat%d(1,1,1,1))
    n_p(2) = C_LOC(pncp%d(1,1,1,1)); n_t(2) = C_LOC(pnct%d(1,1,1,1))
    n_p(3) = C_LOC(pnpp%d(1,1,1,1)); n_t(3) = C_LOC(pnpt%d(1,1,1,1))
    m_p(1) = C_LOC(pmap%d(1,1,1,1)); m_t(1) = C_LOC(pmat%d(1,1,1,1))
    m_p(2) = C_LOC(pmcp%d(1,1,1,1)); m_t(2) = C_LOC(pmct%d(1,1,1,1))
    m_p(3) = C_LOC(pmpp%d(1,1,1,1)); m_t(3) = C_LOC(pmpt%d(1,1,1,1))
    crsi = C_NULL_PTR
    IF (level == 5) THEN
       CALL Diag%getData(1, prsi, name="rsi")
       crsi = C_LOC(prsi%d(1,1,1))
       CALL Prog%getData(1, pnip, name="nice"); CALL Prog%getData(2, pnit, name="nice")
       CALL Prog%getData(1, pmip, name="mice"); CALL Prog%getData(2, pmit, name="mice")
       n_p(4) = C_LOC(pnip%d(1,1,1,1)); n_t(4) = C_LOC(pnit%d(1,1,1,1))
       m_p(4) = C_LOC(pmip%d(1,1,1,1)); m_t(4) = C_LOC(pmit%d(1,1,1,1))
    END IF

    ! process-rate Diag fields: halo values kept, interior overwritten
    nrf = SALSA_NRATE_L4
    IF (level == 5) nrf = SALSA_NRATE_L5
    IF (.NOT. ALLOCATED(rates)) ALLOCATE(rates(nzp,nxp,nyp,nrf))
    DO f = 1, nrf
       CALL Diag%getData(1, fld, name=TRIM(rate_names(f)))
       rates(:,:,:,f) = fld%d(:,:,:)
    END DO

    ! the initialization call (init.f90:101-107) equilibrates the aerosol water
    ! at RH 0.3 before the step (mo_salsa_driver.f90:282-283)
    sflags = SALSA_STEP_TENDENCY
    IF (initialize) sflags = IOR(sflags, SALSA_STEP_INIT)
    rc = salsa_step_host(ctx, C_LOC(ppress%d(1,1,1)), C_LOC(ptk%d(1,1,1)), C_LOC(prs%d(1,1,1)), crsi,   &
                         C_LOC(pdn%d(1,1,1)), C_LOC(prv%d(1,1,1)), C_LOC(prt%d(1,1,1)),                  &
                         n_p, n_t, m_p, m_t, C_LOC(rates(1,1,1,1)), REAL(tstep, C_DOUBLE),               &
                         REAL(time, C_DOUBLE), INT(istp, C_INT32_T), sflags)
    IF (rc /= SALSA_OK) THEN
       WRITE(*,*) 'run_SALSA (libsalsa_amd): ', TRIM(salsa_error_message())
       STOP
    END IF

    DO f = 1, nrf
       CALL Diag%getData(1, fld, name=TRIM(rate_names(f)))
       fld%d(:,:,:) = rates(:,:,:,f)
    END DO

    ! the droplets activated at cloud base (interior cells; the rest stay 0);
    ! pa_vactd's first nwet*ncld slabs, species-major as getMassIndex
    IF (cloud_base) THEN
       rc = hipMemcpy(C_LOC(pa_nactd(1,1,1,1)), d_nactd, 8_C_SIZE_T * n3 * INT(ncld, C_SIZE_T),            &
                      hipMemcpyDeviceToHost)
       IF (rc == 0) rc = hipMemcpy(C_LOC(pa_vactd(1,1,1,1)), d_vactd,                                       &
                                   8_C_SIZE_T * n3 * INT(nwet * ncld, C_SIZE_T), hipMemcpyDeviceToHost)
       IF (rc /= 0) CALL fail('copying pa_nactd / pa_vactd from the device')
    END IF
  END SUBROUTINE run_SALSA

  ! device copies of wp, pa_nactd and pa_vactd, registered with the context
  ! (salsa_set_cloud_base_fields); the kernel writes the interior cells only,
  ! so the rest is zeroed once here
  SUBROUTINE alloc_activation_fields(n3, nwet)
    INTEGER(C_SIZE_T), INTENT(in) :: n3
    INTEGER, INTENT(in) :: nwet
    INTEGER(C_INT) :: rc
    INTEGER(C_SIZE_T) :: bn, bv
    bn = 8_C_SIZE_T * n3 * INT(ncld, C_SIZE_T)
    bv = 8_C_SIZE_T * n3 * INT(nwet * ncld, C_SIZE_T)
    rc = hipMalloc(d_w, 8_C_SIZE_T * n3)
    IF (rc == 0) rc = hipMalloc(d_nactd, bn)
    IF (rc == 0) rc = hipMalloc(d_vactd, bv)
    IF (rc == 0) rc = hipMemset(d_nactd, 0_C_INT, bn)
    IF (rc == 0) rc = hipMemset(d_vactd, 0_C_INT, bv)
    IF (rc /= 0) CALL fail('device memory for the activation fields')
    act_n3 = n3
    rc = salsa_set_cloud_base_fields(ctx, d_w, d_nactd, d_vactd)
    IF (rc /= SALSA_OK) CALL fail('cloud-base activation fields')
  END SUBROUTINE alloc_activation_fields

  SUBROUTINE salsa_amd_finalize()
    INTEGER(C_INT) :: rc
    IF (C_ASSOCIATED(ctx)) CALL salsa_ctx_destroy(ctx)
    ctx = C_NULL_PTR
    IF (C_ASSOCIATED(d_w)) rc = hipFree(d_w)
    IF (C_ASSOCIATED(d_nactd)) rc = hipFree(d_nactd)
    IF (C_ASSOCIATED(d_vactd)) rc = hipFree(d_vactd)
    d_w = C_NULL_PTR; d_nactd = C_NULL_PTR; d_vactd = C_NULL_PTR
    act_n3 = 0
    IF (ALLOCATED(rates)) DEALLOCATE(rates)
  END SUBROUTINE salsa_amd_finalize

  ! Descriptor, switches and context from the SALSA module state
  ! (set_aerobins/set_cloudbins/set_icebins and define_salsa have run).
  SUBROUTINE create_context(level, lcharge, nzp, nxp, nyp)
    INTEGER, INTENT(in) :: level, nzp, nxp, nyp
    LOGICAL, INTENT(in) :: lcharge
    TYPE(salsa_bin_desc) :: desc
    TYPE(salsa_params) :: p
    INTEGER(C_INT32_T) :: ids(7), ndry
    INTEGER :: i, k
    INTEGER(C_INT) :: rc
    CHARACTER(len=16) :: prec

    ndry = INT(spec%getNSpec(type="dry"), C_INT32_T)
    ids = 0
    DO i = 1, ndry
       DO k = 1, 7
          IF (TRIM(spec%names(i)) == TRIM(dry_names(k))) ids(i) = INT(k - 1, C_INT32_T)
       END DO
    END DO
    rc = salsa_bin_desc_init(desc, INT(nbin, C_INT32_T), REAL(reglim, C_DOUBLE),                          &
                             INT(bloPrc%nbins, C_INT32_T), REAL(bloPrc%dlo, C_DOUBLE),                    &
                             REAL(bloPrc%vol_ratio, C_DOUBLE), INT(bloIce%nbins, C_INT32_T),              &
                             REAL(bloIce%dlo, C_DOUBLE), REAL(bloIce%vol_ratio, C_DOUBLE), ndry, ids)
    IF (rc /= SALSA_OK) CALL fail('bin descriptor')
    rc = salsa_params_default(p, INT(level, C_INT32_T))
    IF (rc /= SALSA_OK) CALL fail('parameters')
    p%lscoag = lg(lscoag%state);   p%lscnd = lg(lscnd%state)
    p%lsauto = lg(lsauto%state);   p%lsauto_mode = INT(lsauto%mode, C_INT32_T)
    p%lscgaa = lg(lscgaa); p%lscgcc = lg(lscgcc); p%lscgpp = lg(lscgpp); p%lscgii = lg(lscgii)
    p%lscgca = lg(lscgca); p%lscgpa = lg(lscgpa); p%lscgia = lg(lscgia); p%lscgpc = lg(lscgpc)
    p%lscgic = lg(lscgic); p%lscgip = lg(lscgip)
    p%lscndh2oae = lg(lscndh2oae); p%lscndh2ocl = lg(lscndh2ocl); p%lscndh2oic = lg(lscndh2oic)
    p%lsfreeRH = lg(lsfreeRH%state)
    p%rhlim = REAL(rhlim, C_DOUBLE)
    p%eddy_dis = REAL(eddy_dis, C_DOUBLE)
    p%Eiagg_max = REAL(Eiagg_max, C_DOUBLE); p%Eiagg_min = REAL(Eiagg_min, C_DOUBLE)
    p%ice_alpha = REAL(iceShapeAlpha, C_DOUBLE); p%ice_beta = REAL(iceShapeBeta, C_DOUBLE)
    p%ice_gamma = REAL(iceShapeGamma, C_DOUBLE); p%ice_sigma = REAL(iceShapeSigma, C_DOUBLE)
    p%lcharge = lg(lcharge)
    p%charge_coll_enh = REAL(allSALSA(1,1,1)%chargeCollEnh, C_DOUBLE)
    p%precision = SALSA_FP64
    CALL GET_ENVIRONMENT_VARIABLE('SALSA_AMD_PRECISION', prec)
    IF (TRIM(prec) == 'mixed') p%precision = SALSA_FP32_MIXED
    p%hz = 1; p%hxy = 2                                   ! mo_salsa_driver.f90:180-182
    p%lscoag_mode = INT(lscoag%mode, C_INT32_T)
    p%cgintvl = REAL(cgintvl, C_DOUBLE)
    p%lsdistupdate = MERGE(1_C_INT32_T, 0_C_INT32_T, lsdistupdate)   ! distr_update after condensation
    p%lsactiv = 0_C_INT32_T
    IF (lsactiv%state .AND. lsactiv%mode == 1) p%lsactiv = 1_C_INT32_T   ! interstitial activation
    IF (lsactiv%state .AND. lsactiv%mode == 2) p%lsactiv = 2_C_INT32_T   ! cloud base
    rc = salsa_ctx_create(desc, p, INT(nzp, C_INT32_T), INT(nxp, C_INT32_T), INT(nyp, C_INT32_T), ctx)
    IF (rc /= SALSA_OK) CALL fail('context')
  END SUBROUTINE create_context

  ! Processes that salsa() would run in the same step but this library does not
  SUBROUTINE check_switches()
    IF (lssecice%state) CALL outside('secondary ice production (lssecice)')
    IF (lsactiv%state .AND. lsactiv%mode /= 1 .AND. lsactiv%mode /= 2) CALL outside('cloud activation mode')
    IF (lsicenucl%state .OR. lsicenucl%mode == 2) CALL outside('ice nucleation (lsicenucl)')
    IF (lsicemelt%state) CALL outside('ice melting (lsicemelt)')
    IF (lscndgas) CALL outside('gas-phase condensation (lscndgas)')
    IF (nsnucl > 0) CALL outside('nucleation (nsnucl)')
  END SUBROUTINE check_switches

  SUBROUTINE outside(what)
    CHARACTER(len=*), INTENT(in) :: what
    WRITE(*,*) 'run_SALSA (libsalsa_amd): ', what, ' is switched on but is not part of the GPU step'
    STOP
  END SUBROUTINE outside

  SUBROUTINE fail(what)
    CHARACTER(len=*), INTENT(in) :: what
    WRITE(*,*) 'run_SALSA (libsalsa_amd): ', what, ': ', TRIM(salsa_error_message())
    STOP
  END SUBROUTINE fail

  INTEGER(C_INT32_T) FUNCTION lg(x)
    LOGICAL, INTENT(in) :: x
    lg = 0
    IF (x) lg = 1
  END FUNCTION lg

END MODULE mo_salsa_amd_driver
