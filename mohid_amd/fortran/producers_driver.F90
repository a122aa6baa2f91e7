!------------------------------------------------------------------------------
! producers_driver: a Fortran caller of the upstream producers in
! ModuleAdvectionDiffusionGPU -- the continuity step of ModuleHydrodynamic
! (WaterFluxZGPU) and ModuleGeometry's ComputeVolumes / ComputeAreas
! (ComputeVolumesGPU, ComputeAreasGPU) -- driven from a binary input file so
! that the test suite can compare its results with the oracle.
!
!   producers_driver <input> <output>
!
! Input (stream, native byte order): Size(6), WorkSize(6), variable_grid
! (0/1), DT; then over Size in Fortran order: DUX, DVY, DXX, DYY (2-D real),
! DWZ, DUZ, DVZ, Volume_Z_New, Volume_Z_Old, WaterFlux_X, WaterFlux_Y,
! WaterFlux_Z (3-D real), ComputeFaces3D_U, ComputeFaces3D_V, WaterPoints3D
! (3-D int).  The volumes' and areas' output arrays start from the
! WaterFlux_Z input, so that cells the routines do not write are visible.
! Output: WaterFlux_Z, VolumeZ, VolumeU, VolumeV, AreaU, AreaV.  Exit status
! 0 on success, else a code naming the failing call.
!------------------------------------------------------------------------------
program producers_driver

    use ModuleAdvectionDiffusionGPU
    implicit none

    character(len=1024) :: fin, fout
    integer :: u, STAT, vargrid
    integer :: sz(6), ws(6)
    real(8) :: dt
    real(8), allocatable, dimension(:,:)   :: DUX, DVY, DXX, DYY
    real(8), allocatable, dimension(:,:,:) :: DWZ, DUZ, DVZ, VN, VO, QX, QY, QZ
    real(8), allocatable, dimension(:,:,:) :: VZ, VU, VV, AU, AV
    integer, allocatable, dimension(:,:,:) :: CFU, CFV, WP

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) sz, ws, vargrid, dt
    allocate(DUX(sz(1):sz(2), sz(3):sz(4)), DVY(sz(1):sz(2), sz(3):sz(4)))
    allocate(DXX(sz(1):sz(2), sz(3):sz(4)), DYY(sz(1):sz(2), sz(3):sz(4)))
    allocate(DWZ(sz(1):sz(2), sz(3):sz(4), sz(5):sz(6)))
    allocate(DUZ, DVZ, VN, VO, QX, QY, QZ, VZ, VU, VV, AU, AV, mold=DWZ)
    allocate(CFU(sz(1):sz(2), sz(3):sz(4), sz(5):sz(6)))
    allocate(CFV, WP, mold=CFU)
    read(u) DUX, DVY, DXX, DYY, DWZ, DUZ, DVZ, VN, VO, QX, QY, QZ, CFU, CFV, WP
    close(u)
    VZ = QZ; VU = QZ; VV = QZ; AU = QZ; AV = QZ

    if (vargrid /= 0) then
        call WaterFluxZGPU(sz, ws, QX, QY, QZ, CFU, CFV, WP, dt, Volume_Z_New = VN,       &
                           Volume_Z_Old = VO, STAT = STAT)
    else
        call WaterFluxZGPU(sz, ws, QX, QY, QZ, CFU, CFV, WP, dt, STAT = STAT)
    endif
    if (STAT /= MGPU_SUCCESS) error stop 2
    call ComputeVolumesGPU(sz, ws, DWZ, DUX, DVY, VZ, VU, VV, STAT = STAT)
    if (STAT /= MGPU_SUCCESS) error stop 3
    call ComputeAreasGPU(sz, ws, DUZ, DVZ, DXX, DYY, AU, AV, STAT = STAT)
    if (STAT /= MGPU_SUCCESS) error stop 4

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) QZ, VZ, VU, VV, AU, AV
    close(u)

end program producers_driver
