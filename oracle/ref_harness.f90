! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Box-model driver for the compiled UCLALES-SALSA reference (built from the
! reference sources by oracle/Makefile into oracle/_ref/, never shipped).
! It configures mo_submctl the way define_salsa/NAMELIST would, calls the
! reference's own salsa_initialize, then for every cell in the input file
! packs the Section array as run_SALSA does (mo_salsa_driver.f90:230-279,
! SALSA units already) and calls the reference's coagulation and
! condensation (the enabled branch of salsa(), mo_salsa.f90:92-107).
!
! Modes: 1 = full step(s); 2 = dump the coagulation kernel matrices built
! by update_coagulation_kernels for the input state; 3 = the per-bin ice
! fall-speed chain of DepositionFast (flag 4, mcrp.f90:1146-1176): for every
! ice bin of the input (T in slot 1, air density in slot 2, numbers and
! LES mass mixing ratios in the bin arrays) the reference's calcDiamLES
! (spherical and maximum dimension), getShapeCoefficients and terminal_vel,
! with the air viscosity, mean free path and Cunningham factor formed as
! DepositionFast forms them; 4 = cloud-base activation (cloud_activation
! with lsactiv%mode 2, mo_salsa_cloud.f90:161-375, 1199-1378) of the input
! state with T in slot 1, p in slot 2 and the updraft w in slot 4: the
! activated droplets pactd per cloud bin (number, then the 9 volumes).  With sw(14) = 1 the
! steps of mode 1 use lscoag%mode = 2: the first step of each cell is an
! update step (lcgupdt = .TRUE.), the later ones reuse the kernels left in
! the zcc* arrays by it, which is what fetch_coag provides for a single cell
! (mo_salsa_driver.f90:294-297, :420-422; mo_salsa_dynamics.f90:146-147).
! With sw(15) = 1 each step ends with distr_update (mo_salsa.f90:148-149).
! With lsauto%mode = 2 autoconv2 follows condensation (mo_salsa.f90:110-112);
! with dh(12) = 1 interstitial cloud activation (lsactiv%mode 1) follows it
! (mo_salsa.f90:115-119).  With dh(13) = 1 the first step is an initialization
! call: run_SALSA(initialize=.TRUE.) runs equilibration(init_rh = 0.3,
! init = .TRUE.) on the packed cell before salsa() (mo_salsa_driver.f90:
! 282-283, init_rh :39).  Its initial wet-diameter guess there is whatever
! Section%dwet holds (run_SALSA leaves updateDiameter commented out at :270, so
! the value comes from the cell processed before); this harness gives it the
! cell's own updateDiameter(limit, "all") value instead, which is defined per
! cell.
!
! Usage: ref_harness <input.bin> <output.bin>
PROGRAM ref_harness
  USE mo_submctl, ONLY : nbin, reglim, bloPrc, bloIce, nspec_dry, listspec, lscoag, lscnd, &
                         lsauto, eddy_dis, lscgaa, lscgcc, lscgpp, lscgii, lscgca, lscgpa, &
                         lscgia, lscgpc, lscgic, lscgip, lscndh2oae, lscndh2ocl, lscndh2oic, &
                         lscndgas, nsnucl, lsdistupdate, ntotal, nprc, precpbins, nbins, ncld, nice, &
                         lcgupdt, lsactiv
  USE mo_salsa_init, ONLY : salsa_initialize, associate_master_switches, setDefaultBinLayouts
  USE mo_salsa_types, ONLY : allSALSA, rateDiag, zccaa, zcccc, zccpp, zccii, zccca, &
                             zccpa, zccia, zccpc, zccic, zccip, iaero, icloud, iprecp, iice
  USE mo_salsa_dynamics, ONLY : coagulation, condensation
  USE mo_salsa_update, ONLY : distr_update
  USE mo_salsa_cloud, ONLY : autoconv2, cloud_activation
  USE mo_salsa_properties, ONLY : equilibration
  USE classSection, ONLY : Section
  USE mo_salsa_coagulation_kernels, ONLY : update_coagulation_kernels
  USE classBinLayout, ONLY : BinLayout
  USE classProcessRates, ONLY : Rates
  USE mo_particle_external_properties, ONLY : calcDiamLES, terminal_vel
  USE mo_ice_shape, ONLY : t_shape_coeffs, getShapeCoefficients
  USE mo_submctl, ONLY : spec, pi6
  IMPLICIT NONE

  INTEGER, PARAMETER :: MAGIC = 370010
  CHARACTER(len=512) :: fin, fout
  INTEGER :: hdr(16), ids(7), sw(16)
  REAL :: dh(16)
  INTEGER :: ncell, nstep, nt, level, mode, lch
  INTEGER :: ic, ist, nb, k, i, j, ntot_in
  REAL, ALLOCATABLE :: cellin(:), cellout(:), kmat(:,:)
  REAL :: T(1,1), P(1,1), RS(1,1), RSI(1,1), RV(1,1), dt
  REAL :: g1(1,1), g2(1,1), g3(1,1), g4(1,1), g5(1,1), W(1,1), RH0(1,1)
  LOGICAL :: linit
  TYPE(Section), ALLOCATABLE :: pactd(:,:,:)
  INTEGER :: pbl(1)
  INTEGER(8) :: c0, c1, crate
  REAL :: tcomp
  CHARACTER(len=3), PARAMETER :: names(7) = ['SO4','OC ','BC ','DU ','SS ','NO ','NH ']
  LOGICAL :: lcharge
  ! mode 3 (DepositionFast's constants: defs.f90:45-46, mcrp.f90:1095-1099)
  REAL, PARAMETER :: pi_defs = 3.14159265358979323846264338327, kb_defs = 1.3807e-23
  REAL :: tk1, adn1, avis, va, lambda, zpn1, pm(9), dwet1, dnsp1, zdn1, kn1, gg1, vc1
  TYPE(t_shape_coeffs) :: shp
  INTEGER :: nsi

  CALL get_command_argument(1, fin)
  CALL get_command_argument(2, fout)
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) hdr
  READ(10) ids
  READ(10) dh
  READ(10) sw
  IF (hdr(1) /= MAGIC) STOP 'bad magic'

  ! ---- configuration as define_salsa would leave it (mo_salsa_init.f90:465-637)
  CALL associate_master_switches()
  CALL setDefaultBinLayouts()
  nbin(1) = hdr(3); nbin(2) = hdr(4)
  reglim(1:4) = dh(1:4)
  bloPrc = BinLayout(hdr(5), dh(5), dh(6))
  bloIce = BinLayout(hdr(6), dh(7), dh(8))
  nspec_dry = hdr(7)
  listspec(:) = '   '
  DO k = 1, nspec_dry
     listspec(k) = names(ids(k)+1)
  END DO
  level = hdr(8)
  lscoag%switch = (hdr(9) /= 0);  lscoag%state = lscoag%switch;  lscoag%mode = 1
  lscnd%switch = (hdr(10) /= 0);  lscnd%state = lscnd%switch
  lsauto%switch = (hdr(11) /= 0); lsauto%state = lsauto%switch; lsauto%mode = hdr(12)
  lsactiv%switch = (dh(12) > 0.5); lsactiv%state = lsactiv%switch; lsactiv%mode = 1
  lch = hdr(13)
  lcharge = (lch /= 0)
  nstep = hdr(14)
  ncell = hdr(15)
  ntot_in = hdr(16)
  mode = sw(16)
  linit = (dh(13) > 0.5)
  dt = dh(9)
  eddy_dis = dh(11)
  lscgaa = sw(1) /= 0; lscgcc = sw(2) /= 0; lscgpp = sw(3) /= 0; lscgii = sw(4) /= 0
  lscgca = sw(5) /= 0; lscgpa = sw(6) /= 0; lscgia = sw(7) /= 0; lscgpc = sw(8) /= 0
  lscgic = sw(9) /= 0; lscgip = sw(10) /= 0
  lscndh2oae = sw(11) /= 0; lscndh2ocl = sw(12) /= 0; lscndh2oic = sw(13) /= 0
  lscndgas = .FALSE.; nsnucl = 0
  lsdistupdate = .FALSE.
  IF (level == 4) THEN            ! define_salsa, mo_salsa_init.f90:615-635
     lscgia = .FALSE.; lscgic = .FALSE.; lscgii = .FALSE.; lscgip = .FALSE.
     lscndh2oic = .FALSE.
  END IF

  CALL salsa_initialize(1000., dh(10))
  nt = ntotal
  IF (nt /= ntot_in) STOP 'ntotal mismatch'
  IF (lcharge) allSALSA(:,:,:)%chargeTime = 1000.

  ALLOCATE(cellin(5 + nt + 9*nt), pactd(1,1,ncld))
  W = 0.
  IF (mode == 1) THEN
     ALLOCATE(cellout(1 + nt + 9*nt + 20*10))
  ELSE IF (mode == 3) THEN
     ALLOCATE(cellout(10*nice))
  ELSE IF (mode == 4) THEN
     ALLOCATE(cellout(10*ncld))
  ELSE
     ALLOCATE(cellout(nt*nt), kmat(nt,nt))
  END IF

  OPEN(11, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  ! bin table for layout pinning
  WRITE(11) nt, nprc
  WRITE(11) (allSALSA(1,1,nb)%vlolim, nb=1,nt)
  WRITE(11) (allSALSA(1,1,nb)%vhilim, nb=1,nt)
  WRITE(11) (allSALSA(1,1,nb)%dmid, nb=1,nt)
  WRITE(11) precpbins(1:nprc)

  tcomp = 0.
  DO ic = 1, ncell
     READ(10) cellin
     T(1,1) = cellin(1); P(1,1) = cellin(2); RS(1,1) = cellin(3)
     RSI(1,1) = cellin(4); RV(1,1) = cellin(5)
     DO nb = 1, nt
        allSALSA(1,1,nb)%numc = cellin(5 + nb)
        DO k = 1, 9
           allSALSA(1,1,nb)%volc(k) = cellin(5 + nt + (nb-1)*9 + k)
        END DO
        CALL allSALSA(1,1,nb)%updateRhomean()     ! mo_salsa_driver.f90:272
     END DO
     CALL rateDiag%Reset()

     IF (mode == 1) THEN
        CALL SYSTEM_CLOCK(c0, crate)
        DO ist = 1, nstep
           IF (ist > 1) CALL rateDiag%Reset()
           IF (sw(14) == 1) THEN
              lscoag%mode = 2
              lcgupdt = (ist == 1)
           END IF
           IF (linit .AND. ist == 1) THEN                ! mo_salsa_driver.f90:282-283
              DO nb = 1, nt
                 CALL allSALSA(1,1,nb)%updateDiameter(limit=.TRUE., type="all")
              END DO
              RH0 = 0.3
              CALL equilibration(1, 1, 1, RH0, T, .TRUE.)
           END IF
           IF (lscoag%state) CALL coagulation(1, 1, 1, dt, T, P, lcharge)
           g1 = 0.; g2 = 0.; g3 = 0.; g4 = 0.; g5 = 0.; pbl = 1
           IF (lscnd%state) CALL condensation(1, 1, 1, 1, g1, g2, g3, g4, g5, RV, RS, RSI, &
                                              T, P, dt, pbl)
           IF (lsauto%state .AND. lsauto%mode == 2) CALL autoconv2(1, 1, 1, dt)   ! mo_salsa.f90:110-112
           IF (lsactiv%state) CALL cloud_activation(1, 1, 1, T, P, RV, RS, W, pactd)   ! :115-119
           IF (sw(15) == 1) CALL distr_update(1, 1, level)      ! mo_salsa.f90:148-149
        END DO
        CALL SYSTEM_CLOCK(c1)
        tcomp = tcomp + REAL(c1 - c0) / REAL(crate)
        cellout(1) = RV(1,1)
        DO nb = 1, nt
           cellout(1 + nb) = allSALSA(1,1,nb)%numc
           DO k = 1, 9
              cellout(1 + nt + (nb-1)*9 + k) = allSALSA(1,1,nb)%volc(k)
           END DO
        END DO
        DO k = 1, 20
           cellout(1 + nt + 9*nt + (k-1)*10 + 1:1 + nt + 9*nt + (k-1)*10 + 9) = Rates(k)%volc(1:9)
           cellout(1 + nt + 9*nt + (k-1)*10 + 10) = Rates(k)%numc
        END DO
     ELSE IF (mode == 4) THEN
        lsactiv%mode = 2
        W(1,1) = cellin(4)
        CALL cloud_activation(1, 1, 1, T, P, RV, RS, W, pactd)
        DO i = 1, ncld
           cellout((i-1)*10 + 1) = pactd(1,1,i)%numc
           cellout((i-1)*10 + 2:(i-1)*10 + 10) = pactd(1,1,i)%volc(1:9)
        END DO
     ELSE IF (mode == 3) THEN
        tk1 = cellin(1); adn1 = cellin(2)
        avis = 1.8325e-5*(416.16/(tk1+120.0))*(tk1/296.16)**1.5
        va = sqrt(8.*kb_defs*tk1/(pi_defs*4.8096e-26))
        lambda = 2.*avis/(adn1*va)
        nsi = spec%getNSpec(type='total')
        cellout = 0.
        DO i = 1, nice
           nb = iice + i - 1
           zpn1 = cellin(5 + nb)
           pm(1:nsi) = cellin(5 + nt + (nb-1)*9 + 1:5 + nt + (nb-1)*9 + nsi)
           dwet1 = calcDiamLES(nsi, zpn1, pm(1:nsi), 4, sph=.TRUE.)
           dnsp1 = calcDiamLES(nsi, zpn1, pm(1:nsi), 4, sph=.FALSE.)
           CALL getShapeCoefficients(shp, SUM(pm(1:nsi-1)), pm(nsi), zpn1)
           zdn1 = SUM(pm(1:nsi))/zpn1/(pi6*dwet1**3)
           kn1 = 2.*lambda/dwet1
           gg1 = 1. + kn1*(1.249 + 0.42*exp(-0.87/kn1))
           vc1 = terminal_vel(dwet1, zdn1, adn1, avis, gg1, 4, shp, dnsp1)
           cellout((i-1)*10 + 1:(i-1)*10 + 10) = [dwet1, dnsp1, shp%alpha, shp%beta, shp%gamma, shp%sigma, &
                                                   zdn1, gg1, vc1, avis]
        END DO
     ELSE
        DO nb = 1, nt
           CALL allSALSA(1,1,nb)%updateDiameter(limit=.TRUE., type="all")
        END DO
        CALL update_coagulation_kernels(1, 1, P, T, lcharge)
        kmat = 0.
        IF (lscgaa) CALL put_self(zccaa, iaero, nbins)
        IF (lscgcc) CALL put_self(zcccc, icloud, ncld)
        IF (lscgpp) CALL put_self(zccpp, iprecp, nprc)
        IF (lscgii) CALL put_self(zccii, iice, nice)
        IF (lscgca) CALL put_cross(zccca, iaero, nbins, icloud, ncld)
        IF (lscgpa) CALL put_cross(zccpa, iaero, nbins, iprecp, nprc)
        IF (lscgia) CALL put_cross(zccia, iaero, nbins, iice, nice)
        IF (lscgpc) CALL put_cross(zccpc, icloud, ncld, iprecp, nprc)
        IF (lscgic) CALL put_cross(zccic, icloud, ncld, iice, nice)
        IF (lscgip) CALL put_cross(zccip, iprecp, nprc, iice, nice)
        DO j = 1, nt
           DO i = 1, nt
              cellout((j-1)*nt + i) = kmat(i,j)
           END DO
        END DO
     END IF
     WRITE(11) cellout
  END DO
  CLOSE(10)
  CLOSE(11)
  ! seconds spent in coagulation + condensation (file I/O excluded)
  WRITE(*,'(A,ES20.12)') 'compute_s ', tcomp

CONTAINS

  SUBROUTINE put_self(z, off, n)
    REAL, INTENT(in) :: z(:,:,:,:)
    INTEGER, INTENT(in) :: off, n
    INTEGER :: a, b
    DO b = 1, n
       DO a = 1, n
          kmat(off + a - 1, off + b - 1) = z(1,1,a,b)
       END DO
    END DO
  END SUBROUTINE put_self

  ! zccip is allocated (nice,nprc) but filled and read as (nprc,nice) through
  ! explicit-shape dummies; view the raw storage the same way here.
  SUBROUTINE put_cross(z, off1, n1, off2, n2)
    REAL, INTENT(in) :: z(n1*n2)
    INTEGER, INTENT(in) :: off1, n1, off2, n2
    INTEGER :: a, b
    DO b = 1, n2
       DO a = 1, n1
          kmat(off1 + a - 1, off2 + b - 1) = z((b-1)*n1 + a)
          kmat(off2 + b - 1, off1 + a - 1) = z((b-1)*n1 + a)
       END DO
    END DO
  END SUBROUTINE put_cross

END PROGRAM ref_harness
