! dropin_drive.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Calls the drop-in run_SALSA of mo_salsa_amd_driver the way the LES does:
! the reference's own salsa_initialize configures the SALSA module state
! (bins, species, switches), the Diag and Prog FieldArrays are built with the
! field names initialize_salsa_driver reads (mo_salsa_driver.f90:481-561),
! and run_SALSA is called with the dummy list of mo_salsa_driver.f90:112-133:
! first as the initialization call of init.f90:101-107 (initialize = .TRUE.),
! then as the time-step call of step.f90:317-328, each on zeroed tendencies.
! Built by oracle/Makefile (target dropin) against the reference's module
! files; run by tests/test_gpu_dropin.py.
!
! Usage: dropin_drive <input.bin> <output.bin>
!   input : int32 hdr(12) = magic, nzp, nxp, nyp, level, nbin(1:2), nprc, nice,
!           nspec_dry, istp, activation mode (0 off, 2 cloud base); int32
!           ids(7); real(8) dh(8) = reglim(1:4), prc_dlo, prc_vratio, ice_dlo,
!           ice_vratio; real(8) tstep, time; fields press, temp, rsl, rsi, dn,
!           rp, then per category the number field (nb) and the mass field
!           (nsw*nb), each (nzp,nxp,nyp,.); with cloud-base activation the
!           vertical velocity wp (nzp,nxp,nyp)
!   output: for each of the two calls, the tendencies of rp, then per
!           category of the number and mass fields, then the rate fields;
!           with cloud-base activation then pa_nactd and pa_vactd
PROGRAM dropin_drive
  USE mo_submctl, ONLY : nbin, reglim, bloPrc, bloIce, nspec_dry, listspec, lscoag, lscnd, lsauto, lsactiv, &
                         lsicenucl, lsicemelt, lssecice, lscndgas, nsnucl, lsdistupdate, lscgia, lscgic,    &
                         lscgii, lscgip, lscndh2oic, ncld
  USE mo_salsa_init, ONLY : salsa_initialize, associate_master_switches, setDefaultBinLayouts
  USE classBinLayout, ONLY : BinLayout
  USE classFieldArray, ONLY : FieldArray
  USE mo_structured_datatypes, ONLY : FloatArray3d, FloatArray4d
  USE mo_salsa_amd_driver, ONLY : run_SALSA, salsa_amd_finalize
  IMPLICIT NONE

  INTEGER, PARAMETER :: MAGIC = 370020
  CHARACTER(len=3), PARAMETER :: names(7) = ['SO4','OC ','BC ','DU ','SS ','NO ','NH ']
  CHARACTER(len=16), PARAMETER :: diag_names(5) = [CHARACTER(len=16) :: 'press', 'temp', 'rsl', 'rsi', 'dn']
  CHARACTER(len=16), PARAMETER :: rate_names(17) = [CHARACTER(len=16) ::                                  &
       's_m_autoc', 's_m_autoc50', 's_m_autoc80', 's_m_accr', 's_m_accr50', 's_m_accr80',                 &
       's_m_ACcoll_dry', 's_m_APcoll_dry', 's_n_activ', 's_m_conda', 's_m_condc', 's_m_condp',             &
       's_m_AIcoll_dry', 's_n_icehom', 's_n_icedep', 's_n_iceimm', 's_m_condi']
  CHARACTER(len=8), PARAMETER :: ncat_names(4) = [CHARACTER(len=8) :: 'naero', 'ncloud', 'nprecp', 'nice']
  CHARACTER(len=8), PARAMETER :: mcat_names(4) = [CHARACTER(len=8) :: 'maero', 'mcloud', 'mprecp', 'mice']

  CHARACTER(len=512) :: fin, fout
  INTEGER :: hdr(12), ids(7)
  REAL :: dh(8), tstep, time
  INTEGER :: nzp, nxp, nyp, level, ncat, nwet, c, k, icall, nrf
  INTEGER :: nb(4), nsw(4)
  TYPE(FieldArray) :: Diag, Prog
  TYPE(FloatArray3d), TARGET :: f3(5 + 17), rp_p, rp_t
  TYPE(FloatArray4d), TARGET :: np4(4), nt4(4), mp4(4), mt4(4)
  REAL, ALLOCATABLE, TARGET :: d3(:,:,:,:), rpp(:,:,:), rpt(:,:,:)
  REAL, ALLOCATABLE, TARGET :: nP(:,:,:,:,:), nT(:,:,:,:,:), mP(:,:,:,:,:), mT(:,:,:,:,:)
  REAL, ALLOCATABLE :: wp(:,:,:), nactd(:,:,:,:), vactd(:,:,:,:)
  CLASS(*), POINTER :: pp, pt
  INTEGER :: maxrow

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) hdr
  READ(10) ids
  READ(10) dh
  READ(10) tstep, time
  IF (hdr(1) /= MAGIC) STOP 'bad magic'
  nzp = hdr(2); nxp = hdr(3); nyp = hdr(4); level = hdr(5)

  ! ---- SALSA configuration as define_salsa would leave it (mo_salsa_init.f90:465-637)
  CALL associate_master_switches()
  CALL setDefaultBinLayouts()
  nbin(1) = hdr(6); nbin(2) = hdr(7)
  reglim(1:4) = dh(1:4)
  bloPrc = BinLayout(hdr(8), dh(5), dh(6))
  bloIce = BinLayout(hdr(9), dh(7), dh(8))
  nspec_dry = hdr(10)
  listspec(:) = '   '
  DO k = 1, nspec_dry
     listspec(k) = names(ids(k)+1)
  END DO
  lscoag%switch = .TRUE.; lscoag%state = .TRUE.; lscoag%mode = 1
  lscnd%switch = .TRUE.;  lscnd%state = .TRUE.
  lsauto%switch = .TRUE.; lsauto%state = .TRUE.; lsauto%mode = 1
  lsactiv%state = .FALSE.; lsicenucl%state = .FALSE.; lsicenucl%mode = 1
  IF (hdr(12) == 2) THEN           ! cloud-base activation (mo_salsa.f90:115-119)
     lsactiv%switch = .TRUE.; lsactiv%state = .TRUE.; lsactiv%mode = 2
  END IF
  lsicemelt%state = .FALSE.; lssecice%state = .FALSE.
  lscndgas = .FALSE.; nsnucl = 0
  lsdistupdate = .FALSE.
  IF (level == 4) THEN            ! define_salsa, mo_salsa_init.f90:615-635
     lscgia = .FALSE.; lscgic = .FALSE.; lscgii = .FALSE.; lscgip = .FALSE.
     lscndh2oic = .FALSE.
  END IF
  CALL salsa_initialize(1000., 0.)

  ! ---- LES fields
  ncat = 3
  IF (level == 5) ncat = 4
  nwet = nspec_dry + 1
  nb(1) = 2*nbin(2) + nbin(1); nb(2) = ncld; nb(3) = bloPrc%nbins; nb(4) = bloIce%nbins
  nsw(1:3) = nwet; nsw(4) = nwet + 1
  maxrow = MAXVAL(nb(1:ncat) * nsw(1:ncat))
  ALLOCATE(d3(nzp,nxp,nyp,5+17), rpp(nzp,nxp,nyp), rpt(nzp,nxp,nyp))
  ALLOCATE(nP(nzp,nxp,nyp,MAXVAL(nb),4), nT(nzp,nxp,nyp,MAXVAL(nb),4))
  ALLOCATE(mP(nzp,nxp,nyp,maxrow,4), mT(nzp,nxp,nyp,maxrow,4))
  ALLOCATE(wp(nzp,nxp,nyp), nactd(nzp,nxp,nyp,ncld), vactd(nzp,nxp,nyp,nwet*ncld))
  d3 = 0.; wp = 0.
  DO k = 1, 5
     READ(10) d3(:,:,:,k)
  END DO
  READ(10) rpp
  DO c = 1, ncat
     READ(10) nP(:,:,:,1:nb(c),c)
     READ(10) mP(:,:,:,1:nsw(c)*nb(c),c)
  END DO
  IF (hdr(12) == 2) READ(10) wp
  CLOSE(10)

  Diag = FieldArray()
  Prog = FieldArray()
  DO k = 1, 5 + 17
     f3(k) = FloatArray3d(d3(:,:,:,k))
     pp => f3(k)
     IF (k <= 5) THEN
        CALL Diag%newField(TRIM(diag_names(k)), 'x', 'x', 'N/A', .FALSE., pp)
     ELSE
        CALL Diag%newField(TRIM(rate_names(k-5)), 'x', 'x', 'N/A', .FALSE., pp)
     END IF
  END DO
  rp_p = FloatArray3d(rpp)
  rp_t = FloatArray3d(rpt)
  pp => rp_p; pt => rp_t
  CALL Prog%newField('rp', 'x', 'kg/kg', 'N/A', .FALSE., pp, in_t_data=pt)
  DO c = 1, ncat
     np4(c) = FloatArray4d(nP(:,:,:,1:nb(c),c))
     nt4(c) = FloatArray4d(nT(:,:,:,1:nb(c),c))
     mp4(c) = FloatArray4d(mP(:,:,:,1:nsw(c)*nb(c),c))
     mt4(c) = FloatArray4d(mT(:,:,:,1:nsw(c)*nb(c),c))
     pp => np4(c); pt => nt4(c)
     CALL Prog%newField(TRIM(ncat_names(c)), 'x', '#/kg', 'N/A', .FALSE., pp, in_t_data=pt)
     pp => mp4(c); pt => mt4(c)
     CALL Prog%newField(TRIM(mcat_names(c)), 'x', 'kg/kg', 'N/A', .FALSE., pp, in_t_data=pt)
  END DO

  nrf = 12
  IF (level == 5) nrf = 17
  OPEN(11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  DO icall = 1, 2
     rpt = 0.; nT = 0.; mT = 0.
     d3(:,:,:,6:) = 0.
     ! init.f90:101-107 (istp 0, initialize) then step.f90:317-328
     CALL run_SALSA(Diag, Prog, nzp, nxp, nyp, nwet, wp, nactd, vactd, tstep, time, hdr(11), level, &
                    .FALSE., icall == 1)
     WRITE(11) rpt
     DO c = 1, ncat
        WRITE(11) nT(:,:,:,1:nb(c),c)
        WRITE(11) mT(:,:,:,1:nsw(c)*nb(c),c)
     END DO
     WRITE(11) d3(:,:,:,6:5+nrf)
     IF (hdr(12) == 2) THEN
        WRITE(11) nactd
        WRITE(11) vactd
     END IF
  END DO
  CLOSE(11)
  CALL salsa_amd_finalize()
  WRITE(*,'(A)') 'dropin_drive done'
END PROGRAM dropin_drive
