!------------------------------------------------------------------------------
! adgpu_driver: a Fortran caller of ModuleAdvectionDiffusionGPU, the way
! ModuleWaterProperties calls AdvectionDiffusion (ModuleWaterProperties.F90:
! 4833 Start, 14776 per-property step, 27844 Kill), driven from a binary
! input file so that the test suite can compare its result with the oracle.
!
!   adgpu_driver <input> <output>
!
! Input (stream, native byte order): Size(6), WorkSize(6), nsteps,
! iparam(10) = AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV,
!              Upwind2H, Upwind2V, NullDif, Optimize (0/1),
!              AdvectionNudging (0/1), AdvectionNudgingCells,
! rparam(10) = schmidt_H, SchmidtCoef_V, SchmidtBackground_V, VolumeRelMax,
!              DTProp, ImpExp_AdvV, ImpExp_DifV, ImpExp_AdvXX, ImpExp_AdvYY,
!              ImpExp_DifH,
! then the arrays over Size in Fortran order: DUX, DVY, DZX, DZY (2-D real),
! KFloorZ (2-D int), AreaU, AreaV, DZZ, DWZ, Wflux_X, Wflux_Y, Wflux_Z,
! VolumeZOld, VolumeZ (3-D real), OpenPoints3D, LandPoints3D,
! ComputeFacesU3D, ComputeFacesV3D, ComputeFacesW3D (3-D int), Visc_H, Diff_V
! (3-D real), WaterPoints3D (3-D int), PROP (3-D real), then the
! discharges: DischNumber, the number of discharge cells nc, and when
! DischNumber > 0: DischFlow, DischConc, DischConcMF (nc real), DischI,
! DischJ, DischK, DischKmin, DischKmax (nc int), DischVert, IgnoreDisch,
! DischnCells, ByPass (DischNumber int, 0/1 for the logicals).  The
! discharges are set before every step and unset after it, as
! ModuleWaterProperties does around each call (WP:14763, 14830).  Then the
! optional arguments: iopt(7) = CellFluxes, NoAdvFlux, NoDifFlux (0/1),
! NoFlux arrays given, SmallDepths given (0/1), BoundaryCondition (0:
! absent), ReferenceProp given (0/1); DecayTime (real); NoFluxU, NoFluxV,
! NoFluxW (3-D int) when given; SmallDepths (2-D int, 0/1) when given;
! BoundaryPoints2D (2-D int) when BoundaryCondition or ReferenceProp is
! given; ReferenceProp (3-D real) when given.  An optional argument the
! input does not give is left unallocated, which makes it absent.
! Then the run: mode (0: AdvectionDiffusionGPU per property, as
! WaterProperties calls it; 1: the batched transport, TransportStartGPU ..
! TransportRunGPU), ngeo (0, or nsteps: a geometry per step, a moving free
! surface), ntr (properties), precision (the batched transport's); the
! ntr-1 further properties (3-D real); with ngeo, per step AreaU, AreaV, DZZ,
! DWZ, VolumeZOld, VolumeZ (3-D real), set before the step (SetGeometryGPU /
! TransportSetGeometryGPU, and the step's volumes).
! Output: the ntr properties after nsteps steps, then with CellFluxes the six
! flux arrays of the last step (GetFluxesGPU).  Exit status 0 on success,
! else the STAT.
!------------------------------------------------------------------------------
program adgpu_driver

    use ModuleAdvectionDiffusionGPU
    implicit none

    character(len=1024) :: fin, fout
    integer :: u, s, ID, STAT, nsteps
    integer :: sz(6), ws(6), ipar(10)
    real(8) :: rpar(10)
    real(8), allocatable, dimension(:,:)   :: DUX, DVY, DZX, DZY
    integer, allocatable, dimension(:,:)   :: KFloorZ
    real(8), allocatable, dimension(:,:,:) :: AreaU, AreaV, DZZ, DWZ, Qx, Qy, Qz, VolOld, Vol
    real(8), allocatable, dimension(:,:,:) :: Visc, Diff, PROP
    integer, allocatable, dimension(:,:,:) :: OP, LP, CFU, CFV, CFW, WP
    integer :: ndis, ncell
    real(8), allocatable, dimension(:) :: DFlow, DConc, DConcMF
    integer, allocatable, dimension(:) :: DI, DJ, DK, DKmin, DKmax, DVert, DIgn, DnC, DByp
    integer :: iopt(7)
    real(8) :: rdecay
    integer, allocatable, dimension(:,:,:) :: NFU, NFV, NFW
    integer, allocatable, dimension(:,:)   :: SD, BP
    logical, allocatable, dimension(:,:)   :: SmallD
    real(8), allocatable, dimension(:,:,:) :: RefP, POld, AFX, AFY, AFZ, DFX, DFY, DFZ
    integer, allocatable :: BC
    real(8), allocatable :: Decay
    integer :: run(4), mode, ngeo, ntr, prec, t, TID
    real(8), allocatable, dimension(:,:,:,:) :: PR, GAU, GAV, GZZ, GWZ, GVO, GV

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) sz, ws, nsteps, ipar, rpar
    allocate(DUX(sz(1):sz(2), sz(3):sz(4)), DVY(sz(1):sz(2), sz(3):sz(4)))
    allocate(DZX(sz(1):sz(2), sz(3):sz(4)), DZY(sz(1):sz(2), sz(3):sz(4)))
    allocate(KFloorZ(sz(1):sz(2), sz(3):sz(4)))
    allocate(AreaU(sz(1):sz(2), sz(3):sz(4), sz(5):sz(6)), mold=0.0_8)
    allocate(AreaV, DZZ, DWZ, Qx, Qy, Qz, VolOld, Vol, Visc, Diff, PROP, mold=AreaU)
    allocate(OP(sz(1):sz(2), sz(3):sz(4), sz(5):sz(6)))
    allocate(LP, CFU, CFV, CFW, WP, mold=OP)
    read(u) DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ, Qx, Qy, Qz, VolOld, Vol
    read(u) OP, LP, CFU, CFV, CFW, Visc, Diff, WP, PROP
    read(u) ndis, ncell
    allocate(DFlow(max(1, ncell)), DConc(max(1, ncell)), DConcMF(max(1, ncell)))
    allocate(DI(max(1, ncell)), DJ, DK, DKmin, DKmax, mold=DI)
    allocate(DVert(max(1, ndis)), DIgn, DnC, DByp, mold=DVert)
    if (ndis > 0) then
        read(u) DFlow(1:ncell), DConc(1:ncell), DConcMF(1:ncell)
        read(u) DI(1:ncell), DJ(1:ncell), DK(1:ncell), DKmin(1:ncell), DKmax(1:ncell)
        read(u) DVert(1:ndis), DIgn(1:ndis), DnC(1:ndis), DByp(1:ndis)
    endif
    read(u) iopt, rdecay
    if (iopt(4) /= 0) then
        allocate(NFU, NFV, NFW, mold=OP)
        read(u) NFU, NFV, NFW
    endif
    if (iopt(5) /= 0) then
        allocate(SD(sz(1):sz(2), sz(3):sz(4)), SmallD(sz(1):sz(2), sz(3):sz(4)))
        read(u) SD
        SmallD = SD /= 0
    endif
    if (iopt(6) /= 0 .or. iopt(7) /= 0) then
        allocate(BP(sz(1):sz(2), sz(3):sz(4)))
        read(u) BP
    endif
    if (iopt(6) /= 0) then
        allocate(BC, Decay)
        BC = iopt(6)
        Decay = rdecay
        if (BC == 6) allocate(POld, mold=AreaU)
    endif
    if (iopt(7) /= 0) then
        allocate(RefP, mold=AreaU)
        read(u) RefP
    endif
    read(u) run
    mode = run(1); ngeo = run(2); ntr = run(3); prec = run(4)
    allocate(PR(sz(1):sz(2), sz(3):sz(4), sz(5):sz(6), ntr))
    PR(:, :, :, 1) = PROP
    do t = 2, ntr
        read(u) PR(:, :, :, t)
    enddo
    if (ngeo > 0) then
        allocate(GAU(sz(1):sz(2), sz(3):sz(4), sz(5):sz(6), ngeo))
        allocate(GAV, GZZ, GWZ, GVO, GV, mold=GAU)
        do s = 1, ngeo
            read(u) GAU(:, :, :, s), GAV(:, :, :, s), GZZ(:, :, :, s), GWZ(:, :, :, s)
            read(u) GVO(:, :, :, s), GV(:, :, :, s)
        enddo
    endif
    close(u)

    if (mode == 1) then
        call run_batched()
        call write_out()
        stop
    endif

    call StartAdvectionDiffusionGPU(ID, sz, ws, DUX, DVY, DZX, DZY, KFloorZ,           &
                                    AreaU, AreaV, DZZ, DWZ, 1, STAT = STAT)
    if (STAT /= MGPU_SUCCESS) then
        write(*, *) 'StartAdvectionDiffusionGPU STAT =', STAT
        error stop 2
    endif
    do s = 1, nsteps
      if (ngeo > 0) then
        AreaU = GAU(:, :, :, s); AreaV = GAV(:, :, :, s)
        DZZ = GZZ(:, :, :, s); DWZ = GWZ(:, :, :, s)
        VolOld = GVO(:, :, :, s); Vol = GV(:, :, :, s)
        call SetGeometryGPU(ID, DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ,     &
                            STAT = STAT)
        if (STAT /= MGPU_SUCCESS) then
            write(*, *) 'SetGeometryGPU STAT =', STAT
            error stop 7
        endif
      endif
      do t = 1, ntr
        PROP = PR(:, :, :, t)
        if (ndis > 0) then
            call SetDischargesGPU(ID, DFlow, DConc, DI, DJ, DK, DKmin, DKmax, DVert, ndis,  &
                                  DIgn(1:ndis) /= 0, DnC, DByp(1:ndis) /= 0, DConcMF,      &
                                  STAT = STAT)
            if (STAT /= MGPU_SUCCESS) then
                write(*, *) 'SetDischargesGPU STAT =', STAT
                error stop 4
            endif
        endif
        call AdvectionDiffusionGPU(ID, PROP, rpar(1), rpar(2), rpar(3),                  &
                                   ipar(1), ipar(2), ipar(3), ipar(4),                   &
                                   ipar(5) /= 0, ipar(6) /= 0, rpar(4), ipar(9) /= 0,     &
                                   ipar(10),                                              &
                                   rpar(5), rpar(6), rpar(7), rpar(8), rpar(9), rpar(10), &
                                   ipar(7) /= 0, Qx, Qy, Qz, VolOld, Vol, OP, LP,         &
                                   CFU, CFV, CFW, Visc, Diff,                             &
                                   CellFluxes = iopt(1) /= 0, WaterPoints3D = WP,         &
                                   ReferenceProp = RefP,                                  &
                                   BoundaryCondition = BC, DecayTime = Decay,             &
                                   NumericStability = .false., PROPOld = POld,            &
                                   SmallDepths = SmallD, NoAdvFlux = iopt(2) /= 0,        &
                                   NoDifFlux = iopt(3) /= 0, NoFluxU = NFU,               &
                                   NoFluxV = NFV, NoFluxW = NFW,                          &
                                   Optimize = ipar(8) /= 0, FirstProperty = t == 1,       &
                                   STAT = STAT, BoundaryPoints2D = BP)
        if (STAT /= MGPU_SUCCESS) then
            write(*, *) 'AdvectionDiffusionGPU step', s, 'STAT =', STAT
            error stop 3
        endif
        if (ndis > 0) then
            call UnSetDischargesGPU(ID, STAT = STAT)
            if (STAT /= MGPU_SUCCESS) error stop 5
        endif
        PR(:, :, :, t) = PROP
      enddo
    enddo
    if (iopt(1) /= 0) then
        allocate(AFX, AFY, AFZ, DFX, DFY, DFZ, mold=AreaU)
        call GetFluxesGPU(ID, AFX, AFY, AFZ, DFX, DFY, DFZ, STAT = STAT)
        if (STAT /= MGPU_SUCCESS) then
            write(*, *) 'GetFluxesGPU STAT =', STAT
            error stop 6
        endif
    endif
    call KillAdvectionDiffusionGPU(ID, STAT = STAT)
    call write_out()

contains

    subroutine write_out()
        open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
        write(u) PR
        if (iopt(1) /= 0) write(u) AFX, AFY, AFZ, DFX, DFY, DFZ
        close(u)
    end subroutine write_out

    ! The batched transport: the properties on the GPU for the whole run,
    ! the hydrodynamic state (and with ngeo the geometry) set before each step.
    subroutine run_batched()
        call TransportStartGPU(TID, sz, ws, ntr, prec, DUX, DVY, DZX, DZY, KFloorZ, AreaU,  &
                               AreaV, DZZ, DWZ, 1, STAT = STAT)
        if (STAT /= MGPU_SUCCESS) then
            write(*, *) 'TransportStartGPU STAT =', STAT
            error stop 12
        endif
        do t = 1, ntr
            call TransportSetPropertyGPU(TID, t, PR(:, :, :, t), STAT = STAT)
            if (STAT /= MGPU_SUCCESS) error stop 13
        enddo
        do s = 1, nsteps
            if (ngeo > 0) then
                AreaU = GAU(:, :, :, s); AreaV = GAV(:, :, :, s)
                DZZ = GZZ(:, :, :, s); DWZ = GWZ(:, :, :, s)
                VolOld = GVO(:, :, :, s); Vol = GV(:, :, :, s)
                call TransportSetGeometryGPU(TID, DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, &
                                             DZZ, DWZ, STAT = STAT)
                if (STAT /= MGPU_SUCCESS) error stop 14
            endif
            if (ngeo > 0 .or. s == 1) then
                call TransportSetHydroGPU(TID, rpar(1), rpar(2), rpar(3), ipar(1), ipar(2),  &
                                          ipar(3), ipar(4), ipar(5) /= 0, ipar(6) /= 0,      &
                                          rpar(4), rpar(5), rpar(6), rpar(7), rpar(8),       &
                                          rpar(9), rpar(10), ipar(7) /= 0, Qx, Qy, Qz,       &
                                          VolOld, Vol, OP, LP, CFU, CFV, CFW, Visc, Diff,    &
                                          WP, SmallDepths = SmallD,                          &
                                          NoDifFlux = iopt(3) /= 0, NoFluxU = NFU,           &
                                          NoFluxV = NFV, NoFluxW = NFW,                      &
                                          Optimize = ipar(8) /= 0, STAT = STAT)
                if (STAT /= MGPU_SUCCESS) then
                    write(*, *) 'TransportSetHydroGPU STAT =', STAT
                    error stop 15
                endif
            endif
            if (ngeo > 0) then
                call TransportRunGPU(TID, 1, STAT = STAT)
            else
                call TransportRunGPU(TID, nsteps, STAT = STAT)
            endif
            if (STAT /= MGPU_SUCCESS) then
                write(*, *) 'TransportRunGPU STAT =', STAT
                error stop 16
            endif
            if (ngeo == 0) exit
        enddo
        do t = 1, ntr
            call TransportGetPropertyGPU(TID, t, PR(:, :, :, t), STAT = STAT)
            if (STAT /= MGPU_SUCCESS) error stop 17
        enddo
        call TransportKillGPU(TID, STAT = STAT)
    end subroutine run_batched

end program adgpu_driver
