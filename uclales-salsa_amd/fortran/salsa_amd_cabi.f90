!> ISO_C_BINDING interface to libsalsa_amd (include/salsa_amd.h).
!>
!> The derived types are BIND(C) twins of salsa_parbin / salsa_bin_desc /
!> salsa_params; their components follow the C declarations one for one, so
!> the companion C processor gives both the same layout and padding.
MODULE mo_salsa_amd_cabi
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER, PARAMETER :: SALSA_MAXBIN = 64, SALSA_MAXSPEC = 9

  ! status codes
  INTEGER(C_INT), PARAMETER :: SALSA_OK = 0, SALSA_ERR_ARG = -1, SALSA_ERR_HIP = -2, &
                               SALSA_ERR_NOCONV = -3, SALSA_ERR_NONFINITE = -4, SALSA_ERR_NOMEM = -5, &
                               SALSA_ERR_DISTR = -6
  ! precision modes and step flags
  INTEGER(C_INT32_T), PARAMETER :: SALSA_FP64 = 0, SALSA_FP32_MIXED = 1
  INTEGER(C_INT32_T), PARAMETER :: SALSA_STEP_TENDENCY = 0, SALSA_STEP_INPLACE = 1, &
                                   SALSA_STEP_NO_COMPACT = 2, SALSA_STEP_INIT = 4
  ! process-rate fields
  INTEGER, PARAMETER :: SALSA_NRATE_L4 = 12, SALSA_NRATE_L5 = 17

  TYPE, BIND(C) :: salsa_parbin
     INTEGER(C_INT32_T) :: cur, par
  END TYPE salsa_parbin

  TYPE, BIND(C) :: salsa_bin_desc
     INTEGER(C_INT32_T) :: nbin(2)
     REAL(C_DOUBLE)     :: reglim(4)
     INTEGER(C_INT32_T) :: nprc
     REAL(C_DOUBLE)     :: prc_dlo, prc_vratio
     INTEGER(C_INT32_T) :: nice
     REAL(C_DOUBLE)     :: ice_dlo, ice_vratio
     INTEGER(C_INT32_T) :: nspec_dry
     INTEGER(C_INT32_T) :: dry_species(7)
     INTEGER(C_INT32_T) :: in1a, fn1a, in2a, fn2a, in2b, fn2b, nbins
     TYPE(salsa_parbin) :: ica, fca, icb, fcb
     INTEGER(C_INT32_T) :: ncld
     INTEGER(C_INT32_T) :: ntotal
     INTEGER(C_INT32_T) :: nspec, iwa, irim
     REAL(C_DOUBLE)     :: rholiq(SALSA_MAXSPEC), rhoice(SALSA_MAXSPEC), mm(SALSA_MAXSPEC), diss(SALSA_MAXSPEC)
     INTEGER(C_INT32_T) :: phase(SALSA_MAXBIN)
     REAL(C_DOUBLE)     :: vlolim(SALSA_MAXBIN), vhilim(SALSA_MAXBIN), dmid(SALSA_MAXBIN)
     REAL(C_DOUBLE)     :: nlim(SALSA_MAXBIN), dlim(SALSA_MAXBIN)
     REAL(C_DOUBLE)     :: precpbins(SALSA_MAXBIN)
  END TYPE salsa_bin_desc

  TYPE, BIND(C) :: salsa_params
     INTEGER(C_INT32_T) :: level
     INTEGER(C_INT32_T) :: lscoag, lscnd
     INTEGER(C_INT32_T) :: lsauto, lsauto_mode
     INTEGER(C_INT32_T) :: lscgaa, lscgcc, lscgpp, lscgii, lscgca, lscgpa, lscgia, lscgpc, lscgic, lscgip
     INTEGER(C_INT32_T) :: lscndh2oae, lscndh2ocl, lscndh2oic
     INTEGER(C_INT32_T) :: lsfreeRH
     REAL(C_DOUBLE)     :: rhlim
     REAL(C_DOUBLE)     :: eddy_dis, Eiagg_max, Eiagg_min
     REAL(C_DOUBLE)     :: ice_alpha, ice_beta, ice_gamma, ice_sigma
     INTEGER(C_INT32_T) :: lcharge
     REAL(C_DOUBLE)     :: charge_coll_enh
     INTEGER(C_INT32_T) :: precision
     INTEGER(C_INT32_T) :: hz, hxy
     INTEGER(C_INT32_T) :: lscoag_mode
     REAL(C_DOUBLE)     :: cgintvl
     INTEGER(C_INT32_T) :: lsdistupdate
     INTEGER(C_INT32_T) :: lsactiv
  END TYPE salsa_params

  INTERFACE
     INTEGER(C_INT) FUNCTION salsa_bin_desc_init(desc, nbin, reglim, nprc, prc_dlo, prc_vratio, &
                                                 nice, ice_dlo, ice_vratio, nspec_dry, dry_species) &
                                                 BIND(C, name="salsa_bin_desc_init")
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, salsa_bin_desc
       TYPE(salsa_bin_desc), INTENT(out) :: desc
       INTEGER(C_INT32_T), INTENT(in)    :: nbin(2)
       REAL(C_DOUBLE), INTENT(in)        :: reglim(4)
       INTEGER(C_INT32_T), VALUE         :: nprc
       REAL(C_DOUBLE), VALUE             :: prc_dlo, prc_vratio
       INTEGER(C_INT32_T), VALUE         :: nice
       REAL(C_DOUBLE), VALUE             :: ice_dlo, ice_vratio
       INTEGER(C_INT32_T), VALUE         :: nspec_dry
       INTEGER(C_INT32_T), INTENT(in)    :: dry_species(7)
     END FUNCTION salsa_bin_desc_init

     INTEGER(C_INT) FUNCTION salsa_params_default(p, level) BIND(C, name="salsa_params_default")
       IMPORT :: C_INT, C_INT32_T, salsa_params
       TYPE(salsa_params), INTENT(out) :: p
       INTEGER(C_INT32_T), VALUE       :: level
     END FUNCTION salsa_params_default

     INTEGER(C_INT) FUNCTION salsa_ctx_create(desc, params, nzp, nxp, nyp, ctx) BIND(C, name="salsa_ctx_create")
       IMPORT :: C_INT, C_INT32_T, C_PTR, salsa_bin_desc, salsa_params
       TYPE(salsa_bin_desc), INTENT(in) :: desc
       TYPE(salsa_params), INTENT(in)   :: params
       INTEGER(C_INT32_T), VALUE        :: nzp, nxp, nyp
       TYPE(C_PTR), INTENT(out)         :: ctx
     END FUNCTION salsa_ctx_create

     SUBROUTINE salsa_ctx_destroy(ctx) BIND(C, name="salsa_ctx_destroy")
       IMPORT :: C_PTR
       TYPE(C_PTR), VALUE :: ctx
     END SUBROUTINE salsa_ctx_destroy

     INTEGER(C_INT) FUNCTION salsa_step_host(ctx, press, temp, rsl, rsi, dn, rv_p, rv_t, n_p, n_t, m_p, m_t, &
                                             rates, tstep, time, istp, flags) BIND(C, name="salsa_step_host")
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE        :: ctx, press, temp, rsl, rsi, dn, rv_p, rv_t
       TYPE(C_PTR), INTENT(in)   :: n_p(4), n_t(4), m_p(4), m_t(4)
       TYPE(C_PTR), VALUE        :: rates
       REAL(C_DOUBLE), VALUE     :: tstep, time
       INTEGER(C_INT32_T), VALUE :: istp, flags
     END FUNCTION salsa_step_host

     ! Forward-mode gradient of the precipitation volume after nsteps in-place
     ! steps (device pointers, e.g. from hipMalloc or OpenMP/OpenACC device data)
     INTEGER(C_INT) FUNCTION salsa_gradient_device(ctx, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, &
                                                   nsteps, djdn, j_cell, djde_cell, totals) &
                                                   BIND(C, name="salsa_gradient_device")
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE        :: ctx, press, temp, rsl, rsi, dn, rv
       TYPE(C_PTR), INTENT(in)   :: n_p(4), m_p(4), djdn(4)
       REAL(C_DOUBLE), VALUE     :: tstep
       INTEGER(C_INT32_T), VALUE :: nsteps
       TYPE(C_PTR), VALUE        :: j_cell, djde_cell
       REAL(C_DOUBLE), INTENT(out) :: totals(2)
     END FUNCTION salsa_gradient_device

     ! The same gradient by the reverse (adjoint) sweep, level 4, with the
     ! states kept every checkpoint_every steps and the stretches between
     ! checkpoints recomputed
     INTEGER(C_INT) FUNCTION salsa_adjoint_device(ctx, press, temp, rsl, rsi, dn, rv, n_p, m_p, tstep, &
                                                  nsteps, checkpoint_every, djdn, j_cell, djde_cell, totals) &
                                                  BIND(C, name="salsa_adjoint_device")
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE        :: ctx, press, temp, rsl, rsi, dn, rv
       TYPE(C_PTR), INTENT(in)   :: n_p(4), m_p(4), djdn(4)
       REAL(C_DOUBLE), VALUE     :: tstep
       INTEGER(C_INT32_T), VALUE :: nsteps, checkpoint_every
       TYPE(C_PTR), VALUE        :: j_cell, djde_cell
       REAL(C_DOUBLE), INTENT(out) :: totals(2)
     END FUNCTION salsa_adjoint_device

     ! The LES side after a step on device: tend_constrain2, update_sclrs,
     ! tend0 and SALSA_diagnostics for the SALSA tracers (dzt on the host)
     INTEGER(C_INT) FUNCTION salsa_les_update_device(ctx, dtlt, dzt, temp, rh, rhi, rp, n_p, n_t, m_p, m_t, &
                                                     charge_time, diag) BIND(C, name="salsa_les_update_device")
       IMPORT :: C_INT, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE         :: ctx
       REAL(C_DOUBLE), VALUE      :: dtlt
       REAL(C_DOUBLE), INTENT(in) :: dzt(*)
       TYPE(C_PTR), VALUE         :: temp, rh, rhi, rp
       TYPE(C_PTR), INTENT(in)    :: n_p(4), n_t(4), m_p(4), m_t(4)
       TYPE(C_PTR), VALUE         :: charge_time, diag
     END FUNCTION salsa_les_update_device

     ! Sedimentation of the liquid categories at level 4 (sedim_SALSA)
     INTEGER(C_INT) FUNCTION salsa_sedim_device(ctx, dtlt, dzt, tk, th, adn, ustar, n_p, m_p, n_t, m_t, tlt, &
                                                rrate, sfcrrate, vt_prc, irate, sfcirate, vt_ice, &
                                                sed_aero, sed_cloud, sed_precp, sed_ice) &
                                                BIND(C, name="salsa_sedim_device")
       IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR
       TYPE(C_PTR), VALUE         :: ctx
       REAL(C_DOUBLE), VALUE      :: dtlt
       REAL(C_DOUBLE), INTENT(in) :: dzt(*)
       TYPE(C_PTR), VALUE         :: tk, th, adn, ustar
       TYPE(C_PTR), INTENT(in)    :: n_p(4), m_p(4), n_t(4), m_t(4)
       TYPE(C_PTR), VALUE         :: tlt, rrate, sfcrrate, vt_prc
       TYPE(C_PTR), VALUE         :: irate, sfcirate, vt_ice        ! level 5; C_NULL_PTR at level 4
       INTEGER(C_INT32_T), VALUE  :: sed_aero, sed_cloud, sed_precp, sed_ice
     END FUNCTION salsa_sedim_device

     INTEGER(C_INT) FUNCTION salsa_set_cloud_base_fields(ctx, w, nactd, vactd) &
                                                         BIND(C, name="salsa_set_cloud_base_fields")
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx, w, nactd, vactd         ! device arrays: w, pa_nactd, pa_vactd
     END FUNCTION salsa_set_cloud_base_fields

     INTEGER(C_INT) FUNCTION salsa_set_les_extra_fields(ctx, gaerop, indefp, nindef) &
                                                        BIND(C, name="salsa_set_les_extra_fields")
       IMPORT :: C_INT, C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE        :: ctx, gaerop, indefp     ! device arrays: a_gaerop, a_indefp (or C_NULL_PTR)
       INTEGER(C_INT32_T), VALUE :: nindef                  ! rows of a_indefp
     END FUNCTION salsa_set_les_extra_fields

     INTEGER(C_INT) FUNCTION salsa_ctx_sync(ctx) BIND(C, name="salsa_ctx_sync")
       IMPORT :: C_INT, C_PTR
       TYPE(C_PTR), VALUE :: ctx
     END FUNCTION salsa_ctx_sync

     TYPE(C_PTR) FUNCTION salsa_last_error() BIND(C, name="salsa_last_error")
       IMPORT :: C_PTR
     END FUNCTION salsa_last_error
  END INTERFACE

CONTAINS

  !> The message of salsa_last_error() as a Fortran string.
  FUNCTION salsa_error_message() RESULT(msg)
    CHARACTER(len=512) :: msg
    CHARACTER(kind=C_CHAR), POINTER :: cs(:)
    INTEGER :: i
    msg = ' '
    CALL C_F_POINTER(salsa_last_error(), cs, [512])
    DO i = 1, 512
       IF (cs(i) == C_NULL_CHAR) EXIT
       msg(i:i) = cs(i)
    END DO
  END FUNCTION salsa_error_message

END MODULE mo_salsa_amd_cabi
