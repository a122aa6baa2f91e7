!------------------------------------------------------------------------------
! ModuleAdvectionDiffusionGPU
!
! Fortran host side of the MI355X tracer-transport library: ISO_C_BINDING
! interfaces to the C ABI of include/mohid_gpu.h and Fortran wrappers that
! keep the argument lists of the reference module
!   StartAdvectionDiffusion  Software/MOHIDBase2/ModuleAdvectionDiffusion.F90:400
!   AdvectionDiffusion       Software/MOHIDBase2/ModuleAdvectionDiffusion.F90:1108
!   KillAdvectionDiffusion   Software/MOHIDBase2/ModuleAdvectionDiffusion.F90:5849
!   THOMASZ_NewType2         Software/MOHIDBase1/ModuleFunctions.F90:4026
!   THOMAS_3D                Software/MOHIDBase1/ModuleFunctions.F90:455, 3667
!   SetDischarges            Software/MOHIDBase2/ModuleAdvectionDiffusion.F90:978
!   UnSetDischarges          Software/MOHIDBase2/ModuleAdvectionDiffusion.F90:1040
! so a caller such as ModuleWaterProperties::Advection_Diffusion_Processes
! (Software/MOHIDWater/ModuleWaterProperties.F90:14776) can switch its call.
!
! Arrays are the caller's own: real(8) (the reference's production build is
! default-real-8, Solutions/mohid-in-linux/compile_mohid.sh:185) and integer
! masks spanning Size, passed by the address of their first element.  STAT
! returns 0 on success or the library's status code: MGPU_ERR_UNSUPPORTED
! (103) asks the caller to keep its CPU path for this call.
!------------------------------------------------------------------------------
module ModuleAdvectionDiffusionGPU

    use, intrinsic :: iso_c_binding
    implicit none
    private

    public :: StartAdvectionDiffusionGPU
    public :: AdvectionDiffusionGPU
    public :: KillAdvectionDiffusionGPU
    public :: THOMASZ_GPU
    public :: THOMAS_3D_GPU
    public :: WaterFluxZGPU
    public :: ComputeVolumesGPU
    public :: ComputeAreasGPU
    public :: SetDischargesGPU, UnSetDischargesGPU
    public :: GetFluxesGPU
    public :: SetGeometryGPU
    public :: TransportStartGPU, TransportSetGeometryGPU, TransportSetHydroGPU
    public :: TransportSetPropertyGPU, TransportGetPropertyGPU, TransportRunGPU
    public :: TransportKillGPU
    public :: MGPU_PREC_FP64, MGPU_PREC_FAST, MGPU_PREC_MIXED
    public :: MGPU_SUCCESS, MGPU_OFF_ERR, MGPU_ERR_OPTION, MGPU_ERR_UNSUPPORTED
    public :: MGPU_ERR_PIVOT, MGPU_ERR_BADARG

    integer, parameter :: MGPU_SUCCESS         = 0
    integer, parameter :: MGPU_OFF_ERR         = 1
    integer, parameter :: MGPU_ERR_BADARG      = 101
    integer, parameter :: MGPU_ERR_OPTION      = 102
    integer, parameter :: MGPU_ERR_UNSUPPORTED = 103
    integer, parameter :: MGPU_ERR_PIVOT       = 107
    integer, parameter :: MGPU_ERR_STATE       = 105

    ! arithmetic of the batched transport (include/mohid_gpu.h)
    integer, parameter :: MGPU_PREC_FP64  = 0
    integer, parameter :: MGPU_PREC_FAST  = 1
    integer, parameter :: MGPU_PREC_MIXED = 2

    ! AdvectionDiffusion boundary-condition ids (ModuleAdvectionDiffusion.F90:152-158)
    integer, parameter :: NullGradient_   = 4
    integer, parameter :: CyclicBoundary_ = 8

    type, bind(c) :: t_mgpu_size3d
        integer(c_int32_t) :: ILB, IUB, JLB, JUB, KLB, KUB
    end type t_mgpu_size3d

    type, bind(c) :: t_mgpu_geometry
        type(c_ptr)        :: DUX, DVY, DZX, DZY
        type(c_ptr)        :: KFloorZ
        type(c_ptr)        :: AreaU, AreaV, DZZ, DWZ
        integer(c_int32_t) :: device_pointers
    end type t_mgpu_geometry

    type, bind(c) :: t_mgpu_transport_opts
        integer(c_int32_t) :: ntracers, precision, docycle_method, kernel
        integer(c_int32_t) :: tracers_per_launch, explicit_tracers_per_lane
        integer(c_int32_t) :: column_tracers_per_lane, refinements
    end type t_mgpu_transport_opts

    type, bind(c) :: t_mgpu_ad_args
        real(c_double)     :: schmidt_H, SchmidtCoef_V, SchmidtBackground_V
        integer(c_int32_t) :: AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV
        integer(c_int32_t) :: Upwind2H, Upwind2V
        real(c_double)     :: VolumeRelMax
        integer(c_int32_t) :: AdvectionNudging, AdvectionNudgingCells
        real(c_double)     :: DTProp
        real(c_double)     :: ImpExp_AdvV, ImpExp_DifV, ImpExp_AdvXX, ImpExp_AdvYY, ImpExp_DifH
        integer(c_int32_t) :: NullDif
        type(c_ptr)        :: Wflux_X, Wflux_Y, Wflux_Z
        type(c_ptr)        :: VolumeZOld, VolumeZ
        type(c_ptr)        :: OpenPoints3D, LandPoints3D
        type(c_ptr)        :: ComputeFacesU3D, ComputeFacesV3D, ComputeFacesW3D
        type(c_ptr)        :: Visc_H, Diff_V
        type(c_ptr)        :: WaterPoints3D
        integer(c_int32_t) :: Optimize, FirstProperty
        integer(c_int32_t) :: CellFluxes, NoAdvFlux
        type(c_ptr)        :: ReferenceProp
        integer(c_int32_t) :: BoundaryCondition
        real(c_double)     :: DecayTime
        type(c_ptr)        :: BoundaryPoints2D
        type(c_ptr)        :: PROPOld
        integer(c_int32_t) :: NoDifFlux
        type(c_ptr)        :: NoFluxU, NoFluxV, NoFluxW
        type(c_ptr)        :: SmallDepths
        integer(c_int32_t) :: device_pointers
    end type t_mgpu_ad_args

    interface
        integer(c_int) function mgpu_ad_create(size, worksize, docycle_method, vertical1d, &
                                               xzflow, geometry, handle)                   &
                       bind(c, name="mgpu_ad_create")
            import :: c_int, c_int32_t, c_ptr, t_mgpu_size3d, t_mgpu_geometry
            type(t_mgpu_size3d),   intent(in)  :: size, worksize
            integer(c_int32_t),    value       :: docycle_method, vertical1d, xzflow
            type(t_mgpu_geometry), intent(in)  :: geometry
            type(c_ptr),           intent(out) :: handle
        end function mgpu_ad_create

        integer(c_int) function mgpu_ad_set_geometry(handle, geometry)                      &
                       bind(c, name="mgpu_ad_set_geometry")
            import :: c_int, c_ptr, t_mgpu_geometry
            type(c_ptr),           value      :: handle
            type(t_mgpu_geometry), intent(in) :: geometry
        end function mgpu_ad_set_geometry

        integer(c_int) function mgpu_transport_create(size, worksize, opts, geometry, handle) &
                       bind(c, name="mgpu_transport_create")
            import :: c_int, c_ptr, t_mgpu_size3d, t_mgpu_transport_opts, t_mgpu_geometry
            type(t_mgpu_size3d),         intent(in)  :: size, worksize
            type(t_mgpu_transport_opts), intent(in)  :: opts
            type(t_mgpu_geometry),       intent(in)  :: geometry
            type(c_ptr),                 intent(out) :: handle
        end function mgpu_transport_create

        integer(c_int) function mgpu_transport_set_geometry(handle, geometry)                &
                       bind(c, name="mgpu_transport_set_geometry")
            import :: c_int, c_ptr, t_mgpu_geometry
            type(c_ptr),           value      :: handle
            type(t_mgpu_geometry), intent(in) :: geometry
        end function mgpu_transport_set_geometry

        integer(c_int) function mgpu_transport_set_hydro(handle, args)                       &
                       bind(c, name="mgpu_transport_set_hydro")
            import :: c_int, c_ptr, t_mgpu_ad_args
            type(c_ptr),          value      :: handle
            type(t_mgpu_ad_args), intent(in) :: args
        end function mgpu_transport_set_hydro

        integer(c_int) function mgpu_transport_set_tracer(handle, index, prop, dev)          &
                       bind(c, name="mgpu_transport_set_tracer")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr),        value :: handle, prop
            integer(c_int32_t), value :: index, dev
        end function mgpu_transport_set_tracer

        integer(c_int) function mgpu_transport_get_tracer(handle, index, prop, dev)          &
                       bind(c, name="mgpu_transport_get_tracer")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr),        value :: handle, prop
            integer(c_int32_t), value :: index, dev
        end function mgpu_transport_get_tracer

        integer(c_int) function mgpu_transport_run(handle, nsteps, use_graph)                &
                       bind(c, name="mgpu_transport_run")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr),        value :: handle
            integer(c_int32_t), value :: nsteps, use_graph
        end function mgpu_transport_run

        integer(c_int) function mgpu_transport_status(handle)                                &
                       bind(c, name="mgpu_transport_status")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function mgpu_transport_status

        integer(c_int) function mgpu_transport_destroy(handle)                               &
                       bind(c, name="mgpu_transport_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function mgpu_transport_destroy

        integer(c_int) function mgpu_ad_step(handle, args, prop) bind(c, name="mgpu_ad_step")
            import :: c_int, c_ptr, t_mgpu_ad_args
            type(c_ptr),          value      :: handle
            type(t_mgpu_ad_args), intent(in) :: args
            type(c_ptr),          value      :: prop
        end function mgpu_ad_step

        integer(c_int) function mgpu_ad_destroy(handle) bind(c, name="mgpu_ad_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function mgpu_ad_destroy

        integer(c_int) function mgpu_thomasz(size, worksize, D, E, F, TI, waterpoints, res, &
                                             dev, stream) bind(c, name="mgpu_thomasz")
            import :: c_int, c_int32_t, c_ptr, t_mgpu_size3d
            type(t_mgpu_size3d), intent(in) :: size, worksize
            type(c_ptr),         value      :: D, E, F, TI, waterpoints, res
            integer(c_int32_t),  value      :: dev
            type(c_ptr),         value      :: stream
        end function mgpu_thomasz

        integer(c_int) function mgpu_thomas3d(size, IJmin, IJmax, JImin, JImax, Kmin, Kmax, &
                                              di, dj, D, E, F, TI, res, dev, stream)       &
                                              bind(c, name="mgpu_thomas3d")
            import :: c_int, c_int32_t, c_ptr, t_mgpu_size3d
            type(t_mgpu_size3d), intent(in) :: size
            integer(c_int32_t),  value      :: IJmin, IJmax, JImin, JImax, Kmin, Kmax, di, dj
            type(c_ptr),         value      :: D, E, F, TI, res
            integer(c_int32_t),  value      :: dev
            type(c_ptr),         value      :: stream
        end function mgpu_thomas3d

        integer(c_int) function mgpu_ad_set_discharges(handle, DischFlow, DischConc, DischI,   &
                DischJ, DischK, DischKmin, DischKmax, DischVert, DischNumber, IgnoreDisch,    &
                DischnCells, ByPass, DischConcMF) bind(c, name="mgpu_ad_set_discharges")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr),        value :: handle
            type(c_ptr),        value :: DischFlow, DischConc, DischI, DischJ, DischK
            type(c_ptr),        value :: DischKmin, DischKmax, DischVert
            integer(c_int32_t), value :: DischNumber
            type(c_ptr),        value :: IgnoreDisch, DischnCells, ByPass, DischConcMF
        end function mgpu_ad_set_discharges

        integer(c_int) function mgpu_ad_unset_discharges(handle)                               &
                bind(c, name="mgpu_ad_unset_discharges")
            import :: c_int, c_ptr
            type(c_ptr), value :: handle
        end function mgpu_ad_unset_discharges

        integer(c_int) function mgpu_ad_get_fluxes(handle, afx, afy, afz, dfx, dfy, dfz, dev) &
                bind(c, name="mgpu_ad_get_fluxes")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: handle, afx, afy, afz, dfx, dfy, dfz
            integer(c_int32_t), value :: dev
        end function mgpu_ad_get_fluxes

        integer(c_int) function mgpu_waterflux_z(size, worksize, vnew, vold, qx, qy, qz, &
                                                 cfu, cfv, wp, dt, vargrid, dev)           &
                bind(c, name="mgpu_waterflux_z")
            import :: c_int, c_int32_t, c_ptr, c_double, t_mgpu_size3d
            type(t_mgpu_size3d), intent(in) :: size, worksize
            type(c_ptr),         value      :: vnew, vold, qx, qy, qz, cfu, cfv, wp
            real(c_double),      value      :: dt
            integer(c_int32_t),  value      :: vargrid, dev
        end function mgpu_waterflux_z

        integer(c_int) function mgpu_compute_volumes(size, worksize, dwz, dux, dvy, vz, vu, &
                                                     vv, vw, dev)                           &
                bind(c, name="mgpu_compute_volumes")
            import :: c_int, c_int32_t, c_ptr, t_mgpu_size3d
            type(t_mgpu_size3d), intent(in) :: size, worksize
            type(c_ptr),         value      :: dwz, dux, dvy, vz, vu, vv, vw
            integer(c_int32_t),  value      :: dev
        end function mgpu_compute_volumes

        integer(c_int) function mgpu_compute_areas(size, worksize, duz, dvz, dxx, dyy, au, av, &
                                                   szz, cu, cxu, cv, cxv, dev)                 &
                bind(c, name="mgpu_compute_areas")
            import :: c_int, c_int32_t, c_ptr, t_mgpu_size3d
            type(t_mgpu_size3d), intent(in) :: size, worksize
            type(c_ptr),         value      :: duz, dvz, dxx, dyy, au, av, szz, cu, cxu, cv, cxv
            integer(c_int32_t),  value      :: dev
        end function mgpu_compute_areas
    end interface

    ! instance table: integer IDs as MOHID's modules hand out (ModuleGlobalData
    ! RegisterNewInstance), each holding one library handle
    integer, parameter :: MAX_INSTANCES = 64
    type(c_ptr), dimension(MAX_INSTANCES), save :: Handles = c_null_ptr
    ! batched transports (the tracer loop of Advection_Diffusion_Processes)
    type(c_ptr), dimension(MAX_INSTANCES), save :: THandles = c_null_ptr

contains

    function to_size3d(b) result(s)
        integer, dimension(6), intent(in) :: b
        type(t_mgpu_size3d) :: s
        s%ILB = b(1); s%IUB = b(2); s%JLB = b(3); s%JUB = b(4); s%KLB = b(5); s%KUB = b(6)
    end function to_size3d

    function b2i(l) result(i)
        logical, intent(in) :: l
        integer(c_int32_t)  :: i
        i = 0
        if (l) i = 1
    end function b2i

    !> StartAdvectionDiffusion plus the geometry AdvectionDiffusion fetches
    !> from HorizontalGrid/Geometry (ModuleAdvectionDiffusion.F90:1353-1401).
    !> Size/WorkSize: ILB, IUB, JLB, JUB, KLB, KUB (T_Size3D order).
    subroutine StartAdvectionDiffusionGPU(ID, Size, WorkSize, DUX, DVY, DZX, DZY, KFloorZ, &
                                          AreaU, AreaV, DZZ, DWZ, Docycle_method,            &
                                          Vertical1D, XZFlow, STAT)
        integer,                           intent(out) :: ID
        integer, dimension(6),             intent(in)  :: Size, WorkSize
        real(8), dimension(:,:),   target, intent(in)  :: DUX, DVY, DZX, DZY
        integer, dimension(:,:),   target, intent(in)  :: KFloorZ
        real(8), dimension(:,:,:), contiguous, target, intent(in)  :: AreaU, AreaV, DZZ, DWZ
        integer,                           intent(in)  :: Docycle_method
        logical, optional,                 intent(in)  :: Vertical1D, XZFlow
        integer, optional,                 intent(out) :: STAT
        type(t_mgpu_geometry) :: g
        type(t_mgpu_size3d)   :: s, w
        type(c_ptr)           :: h
        integer               :: rc, n
        logical               :: v1d, xz

        v1d = .false.; if (present(Vertical1D)) v1d = Vertical1D
        xz  = .false.; if (present(XZFlow))     xz  = XZFlow
        g%DUX = c_loc(DUX); g%DVY = c_loc(DVY); g%DZX = c_loc(DZX); g%DZY = c_loc(DZY)
        g%KFloorZ = c_loc(KFloorZ)
        g%AreaU = c_loc(AreaU); g%AreaV = c_loc(AreaV); g%DZZ = c_loc(DZZ); g%DWZ = c_loc(DWZ)
        g%device_pointers = 0
        s = to_size3d(Size); w = to_size3d(WorkSize)
        ID = 0
        rc = mgpu_ad_create(s, w, int(Docycle_method, c_int32_t), b2i(v1d), b2i(xz), g, h)
        if (rc == MGPU_SUCCESS) then
            do n = 1, MAX_INSTANCES
                if (.not. c_associated(Handles(n))) then
                    Handles(n) = h
                    ID = n
                    exit
                endif
            enddo
            if (ID == 0) then
                rc = mgpu_ad_destroy(h)
                rc = -99
            endif
        endif
        if (present(STAT)) STAT = rc
    end subroutine StartAdvectionDiffusionGPU

    function geometry_of(DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ) result(g)
        real(8), dimension(:,:),   target, intent(in) :: DUX, DVY, DZX, DZY
        integer, dimension(:,:),   target, intent(in) :: KFloorZ
        real(8), dimension(:,:,:), contiguous, target, intent(in) :: AreaU, AreaV, DZZ, DWZ
        type(t_mgpu_geometry) :: g
        g%DUX = c_loc(DUX); g%DVY = c_loc(DVY); g%DZX = c_loc(DZX); g%DZY = c_loc(DZY)
        g%KFloorZ = c_loc(KFloorZ)
        g%AreaU = c_loc(AreaU); g%AreaV = c_loc(AreaV); g%DZZ = c_loc(DZZ); g%DWZ = c_loc(DWZ)
        g%device_pointers = 0
    end function geometry_of

    !> The geometry of the following AdvectionDiffusionGPU calls.  The
    !> reference fetches DZZ, DWZ, AreaU and AreaV from ModuleGeometry inside
    !> every call (GetGeometryDistances / GetGeometryAreas,
    !> ModuleAdvectionDiffusion.F90:1353-1401), and ModuleGeometry recomputes
    !> them every step as the free surface moves (ComputeVolumes /
    !> ComputeAreas, ModuleGeometry.F90:4108-4327); call this once per step,
    !> before the step's first AdvectionDiffusionGPU call, with the arrays
    !> StartAdvectionDiffusionGPU took.
    subroutine SetGeometryGPU(ID, DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ, STAT)
        integer,                           intent(in)  :: ID
        real(8), dimension(:,:),   target, intent(in)  :: DUX, DVY, DZX, DZY
        integer, dimension(:,:),   target, intent(in)  :: KFloorZ
        real(8), dimension(:,:,:), contiguous, target, intent(in)  :: AreaU, AreaV, DZZ, DWZ
        integer, optional,                 intent(out) :: STAT
        type(t_mgpu_geometry) :: g
        integer               :: rc
        if (ID < 1 .or. ID > MAX_INSTANCES) then
            if (present(STAT)) STAT = MGPU_OFF_ERR
            return
        endif
        g = geometry_of(DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ)
        rc = mgpu_ad_set_geometry(Handles(ID), g)
        if (present(STAT)) STAT = rc
    end subroutine SetGeometryGPU

    !> AdvectionDiffusion (ModuleAdvectionDiffusion.F90:1108-1198): PROP is
    !> advanced by DTProp in place.  Arguments keep the reference's names,
    !> order and meaning, so WaterProperties' call (WP:14776-14822) compiles
    !> unchanged against it once renamed and given BoundaryPoints2D, the one
    !> argument added (after STAT): the array the reference fetches from
    !> HorizontalMap inside the call (GetBoundaries, AD:1373).
    !> NumericStability is accepted and ignored: the reference only stores it
    !> (AD:1265-1268).  An absent FirstProperty counts as .true. (nothing is
    !> kept from an earlier call).  Options the GPU path does not offload make
    !> STAT = MGPU_ERR_UNSUPPORTED and leave PROP untouched:
    !>   CellFluxes with implicit horizontal advection, an open boundary with
    !>   one layer and implicit horizontal advection, the Orlanski_ open
    !>   boundary with implicit horizontal advection, and AdvectionNudging
    !>   with Optimize, Vertical1D or implicit horizontal advection.
    !> AdvectionNudging (AD:1989-2068), NoAdvFlux (with NoFluxU/V), CellFluxes,
    !> SmallDepths, NoDifFlux (with NoFluxU/V/W), the discharges of
    !> SetDischargesGPU, the open boundary (ReferenceProp, BoundaryCondition,
    !> DecayTime, AD:5369-5665) and the NullGradient_ / CyclicBoundary_ passes
    !> after the solve (AD:1874-1880) are offloaded.  BoundaryPoints2D is the
    !> array the reference fetches from HorizontalMap (GetBoundaries, AD:1373);
    !> it is needed with ReferenceProp or those two conditions.
    !> FirstProperty is .true. on the first call of every step (WP:14600,
    !> 14821-14826): the library keeps, like the reference, coefficients of
    !> the previous call of the same step.
    subroutine AdvectionDiffusionGPU(ID, PROP, schmidt_H, SchmidtCoef_V, SchmidtBackground_V, &
                                     AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV,  &
                                     Upwind2H, Upwind2V, VolumeRelMax, AdvectionNudging,      &
                                     AdvectionNudgingCells, DTProp, ImpExp_AdvV, ImpExp_DifV, &
                                     ImpExp_AdvXX, ImpExp_AdvYY, ImpExp_DifH, NullDif,        &
                                     Wflux_X, Wflux_Y, Wflux_Z, VolumeZOld, VolumeZ,          &
                                     OpenPoints3D, LandPoints3D, ComputeFacesU3D,             &
                                     ComputeFacesV3D, ComputeFacesW3D, Visc_H, Diff_V,        &
                                     CellFluxes, WaterPoints3D, ReferenceProp,                &
                                     BoundaryCondition, DecayTime, NumericStability, PROPOld, &
                                     SmallDepths, NoAdvFlux, NoDifFlux,                       &
                                     NoFluxU, NoFluxV, NoFluxW,                               &
                                     Optimize, FirstProperty, STAT, BoundaryPoints2D)
        integer,                           intent(in)    :: ID
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: PROP
        real(8),                           intent(in)    :: schmidt_H, SchmidtCoef_V
        real(8),                           intent(in)    :: SchmidtBackground_V
        integer,                           intent(in)    :: AdvMethodH, TVDLimitationH
        integer,                           intent(in)    :: AdvMethodV, TVDLimitationV
        logical,                           intent(in)    :: Upwind2H, Upwind2V
        real(8),                           intent(in)    :: VolumeRelMax
        logical,                           intent(in)    :: AdvectionNudging
        integer,                           intent(in)    :: AdvectionNudgingCells
        real(8),                           intent(in)    :: DTProp
        real(8),                           intent(in)    :: ImpExp_AdvV, ImpExp_DifV
        real(8),                           intent(in)    :: ImpExp_AdvXX, ImpExp_AdvYY
        real(8),                           intent(in)    :: ImpExp_DifH
        logical,                           intent(in)    :: NullDif
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: Wflux_X, Wflux_Y, Wflux_Z
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: VolumeZOld, VolumeZ
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: OpenPoints3D, LandPoints3D
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: ComputeFacesU3D, ComputeFacesV3D
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: ComputeFacesW3D
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: Visc_H, Diff_V
        integer, dimension(:,:,:), optional, contiguous, target, intent(in) :: WaterPoints3D
        logical, optional,                 intent(in)    :: CellFluxes
        logical, optional,                 intent(in)    :: NumericStability
        real(8), dimension(:,:,:), optional, contiguous, target, intent(in) :: ReferenceProp
        integer, optional,                 intent(in)    :: BoundaryCondition
        real(8), optional,                 intent(in)    :: DecayTime
        integer, dimension(:,:), optional, contiguous, target, intent(in) :: BoundaryPoints2D
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: PROPOld
        logical, dimension(:,:), optional, intent(in)    :: SmallDepths
        logical, optional,                 intent(in)    :: NoAdvFlux, NoDifFlux
        integer, dimension(:,:,:), optional, contiguous, target, intent(in) :: NoFluxU, NoFluxV
        integer, dimension(:,:,:), optional, contiguous, target, intent(in) :: NoFluxW
        logical, optional,                 intent(in)    :: Optimize, FirstProperty
        integer, optional,                 intent(out)   :: STAT
        type(t_mgpu_ad_args) :: a
        integer              :: rc
        integer, allocatable, target :: small(:,:)

        if (ID < 1 .or. ID > MAX_INSTANCES) then
            if (present(STAT)) STAT = MGPU_OFF_ERR
            return
        endif
        a%schmidt_H = schmidt_H
        a%SchmidtCoef_V = SchmidtCoef_V
        a%SchmidtBackground_V = SchmidtBackground_V
        a%AdvMethodH = AdvMethodH; a%TVDLimitationH = TVDLimitationH
        a%AdvMethodV = AdvMethodV; a%TVDLimitationV = TVDLimitationV
        a%Upwind2H = b2i(Upwind2H); a%Upwind2V = b2i(Upwind2V)
        a%VolumeRelMax = VolumeRelMax
        a%AdvectionNudging = b2i(AdvectionNudging)
        a%AdvectionNudgingCells = AdvectionNudgingCells
        a%DTProp = DTProp
        a%ImpExp_AdvV = ImpExp_AdvV; a%ImpExp_DifV = ImpExp_DifV
        a%ImpExp_AdvXX = ImpExp_AdvXX; a%ImpExp_AdvYY = ImpExp_AdvYY
        a%ImpExp_DifH = ImpExp_DifH
        a%NullDif = b2i(NullDif)
        a%Wflux_X = c_loc(Wflux_X); a%Wflux_Y = c_loc(Wflux_Y); a%Wflux_Z = c_loc(Wflux_Z)
        a%VolumeZOld = c_loc(VolumeZOld); a%VolumeZ = c_loc(VolumeZ)
        a%OpenPoints3D = c_loc(OpenPoints3D); a%LandPoints3D = c_loc(LandPoints3D)
        a%ComputeFacesU3D = c_loc(ComputeFacesU3D)
        a%ComputeFacesV3D = c_loc(ComputeFacesV3D)
        a%ComputeFacesW3D = c_loc(ComputeFacesW3D)
        a%Visc_H = c_loc(Visc_H); a%Diff_V = c_loc(Diff_V)
        ! the library needs WaterPoints3D (the columns THOMASZ solves, MF:4086)
        if (.not. present(WaterPoints3D)) then
            if (present(STAT)) STAT = MGPU_ERR_BADARG
            return
        endif
        a%WaterPoints3D = c_loc(WaterPoints3D)
        a%Optimize = 0; if (present(Optimize)) a%Optimize = b2i(Optimize)
        ! without FirstProperty nothing is kept from an earlier call: every
        ! coefficient is made again (the safe reading of an absent flag)
        a%FirstProperty = 1; if (present(FirstProperty)) a%FirstProperty = b2i(FirstProperty)
        a%CellFluxes = 0
        if (present(CellFluxes)) a%CellFluxes = b2i(CellFluxes)
        a%ReferenceProp = c_null_ptr
        if (present(ReferenceProp)) a%ReferenceProp = c_loc(ReferenceProp)
        a%BoundaryCondition = 0
        if (present(BoundaryCondition)) a%BoundaryCondition = BoundaryCondition
        a%DecayTime = -9.9e15_8
        if (present(DecayTime)) a%DecayTime = DecayTime
        a%BoundaryPoints2D = c_null_ptr
        if (present(BoundaryPoints2D)) a%BoundaryPoints2D = c_loc(BoundaryPoints2D)
        a%PROPOld = c_null_ptr
        if (present(PROPOld)) a%PROPOld = c_loc(PROPOld)
        a%NoAdvFlux = 0
        if (present(NoAdvFlux)) a%NoAdvFlux = b2i(NoAdvFlux)
        a%NoDifFlux = 0
        if (present(NoDifFlux)) a%NoDifFlux = b2i(NoDifFlux)
        a%NoFluxU = c_null_ptr; a%NoFluxV = c_null_ptr; a%NoFluxW = c_null_ptr
        if (present(NoFluxU)) a%NoFluxU = c_loc(NoFluxU)
        if (present(NoFluxV)) a%NoFluxV = c_loc(NoFluxV)
        if (present(NoFluxW)) a%NoFluxW = c_loc(NoFluxW)
        a%SmallDepths = c_null_ptr
        if (present(SmallDepths)) then
            allocate(small(size(SmallDepths, 1), size(SmallDepths, 2)))
            small = merge(1, 0, SmallDepths)
            a%SmallDepths = c_loc(small)
        endif
        a%device_pointers = 0
        rc = mgpu_ad_step(Handles(ID), a, c_loc(PROP))
        if (present(STAT)) STAT = rc
    end subroutine AdvectionDiffusionGPU

    !> SetDischarges (ModuleAdvectionDiffusion.F90:978-1034) with the reference's
    !> arguments: the discharges of the following AdvectionDiffusionGPU calls
    !> (the arrays are copied; WaterProperties sets them before every call,
    !> WP:14763).
    subroutine SetDischargesGPU(ID, DischFlow, DischConc, DischI, DischJ, DischK,        &
                                DischKmin, DischKmax, DischVert, DischNumber,             &
                                IgnoreDisch, DischnCells, ByPass, DischConcMF, STAT)
        integer,                        intent(in)  :: ID
        real(8), dimension(:), target,  intent(in)  :: DischFlow, DischConc, DischConcMF
        integer, dimension(:), target,  intent(in)  :: DischI, DischJ, DischK
        integer, dimension(:), target,  intent(in)  :: DischKmin, DischKmax, DischVert
        integer,                        intent(in)  :: DischNumber
        logical, dimension(:),          intent(in)  :: IgnoreDisch, ByPass
        integer, dimension(:), target,  intent(in)  :: DischnCells
        integer, optional,              intent(out) :: STAT
        integer, allocatable, target :: ign(:), byp(:)
        integer :: rc
        if (ID < 1 .or. ID > MAX_INSTANCES) then
            if (present(STAT)) STAT = MGPU_OFF_ERR
            return
        endif
        allocate(ign(max(1, size(IgnoreDisch))), byp(max(1, size(ByPass))))
        ign = 0; byp = 0
        ign(1:size(IgnoreDisch)) = merge(1, 0, IgnoreDisch)
        byp(1:size(ByPass)) = merge(1, 0, ByPass)
        rc = mgpu_ad_set_discharges(Handles(ID), c_loc(DischFlow), c_loc(DischConc),          &
                                    c_loc(DischI), c_loc(DischJ), c_loc(DischK),              &
                                    c_loc(DischKmin), c_loc(DischKmax), c_loc(DischVert),     &
                                    int(DischNumber, c_int32_t), c_loc(ign),                  &
                                    c_loc(DischnCells), c_loc(byp), c_loc(DischConcMF))
        if (present(STAT)) STAT = rc
    end subroutine SetDischargesGPU

    !> GetAdvFlux / GetDifFlux (ModuleAdvectionDiffusion.F90:697-860): the cell
    !> fluxes of the last AdvectionDiffusionGPU call made with CellFluxes, into
    !> the caller's arrays over Size (each optional).
    subroutine GetFluxesGPU(ID, AdvFluxX, AdvFluxY, AdvFluxZ, DifFluxX, DifFluxY,       &
                            DifFluxZ, STAT)
        integer,                                  intent(in)  :: ID
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: AdvFluxX
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: AdvFluxY
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: AdvFluxZ
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: DifFluxX
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: DifFluxY
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: DifFluxZ
        integer, optional,                        intent(out) :: STAT
        type(c_ptr) :: p(6)
        integer :: rc
        if (ID < 1 .or. ID > MAX_INSTANCES) then
            if (present(STAT)) STAT = MGPU_OFF_ERR
            return
        endif
        p = c_null_ptr
        if (present(AdvFluxX)) p(1) = c_loc(AdvFluxX)
        if (present(AdvFluxY)) p(2) = c_loc(AdvFluxY)
        if (present(AdvFluxZ)) p(3) = c_loc(AdvFluxZ)
        if (present(DifFluxX)) p(4) = c_loc(DifFluxX)
        if (present(DifFluxY)) p(5) = c_loc(DifFluxY)
        if (present(DifFluxZ)) p(6) = c_loc(DifFluxZ)
        rc = mgpu_ad_get_fluxes(Handles(ID), p(1), p(2), p(3), p(4), p(5), p(6), 0_c_int32_t)
        if (present(STAT)) STAT = rc
    end subroutine GetFluxesGPU

    !> UnSetDischarges (ModuleAdvectionDiffusion.F90:1040-1086).
    subroutine UnSetDischargesGPU(ID, STAT)
        integer,           intent(in)  :: ID
        integer, optional, intent(out) :: STAT
        integer :: rc
        rc = MGPU_OFF_ERR
        if (ID >= 1 .and. ID <= MAX_INSTANCES) rc = mgpu_ad_unset_discharges(Handles(ID))
        if (present(STAT)) STAT = rc
    end subroutine UnSetDischargesGPU

    !> KillAdvectionDiffusion (ModuleAdvectionDiffusion.F90:5849).
    !> The tracer loop of ModuleWaterProperties::Advection_Diffusion_Processes
    !> (WP:14520-15169) for NTracers properties that share one hydrodynamic
    !> state, kept on the GPU across steps: TransportStartGPU (with the
    !> geometry StartAdvectionDiffusionGPU takes), TransportSetPropertyGPU per
    !> property, then per step TransportSetGeometryGPU (a moving vertical
    !> coordinate), TransportSetHydroGPU (the hydrodynamic arguments of
    !> AdvectionDiffusion) and TransportRunGPU; TransportGetPropertyGPU
    !> returns a property's concentration.  Every property of a step gets the
    !> per-property semantics of AdvectionDiffusion, Optimize/FirstProperty as
    !> WP:14576-14598 decide.  Precision: MGPU_PREC_FP64 (the reference's
    !> arithmetic), MGPU_PREC_FAST, MGPU_PREC_MIXED.
    subroutine TransportStartGPU(TID, Size, WorkSize, NTracers, Precision, DUX, DVY, DZX,    &
                                 DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ, Docycle_method, STAT)
        integer,                           intent(out) :: TID
        integer, dimension(6),             intent(in)  :: Size, WorkSize
        integer,                           intent(in)  :: NTracers, Precision
        real(8), dimension(:,:),   target, intent(in)  :: DUX, DVY, DZX, DZY
        integer, dimension(:,:),   target, intent(in)  :: KFloorZ
        real(8), dimension(:,:,:), contiguous, target, intent(in)  :: AreaU, AreaV, DZZ, DWZ
        integer,                           intent(in)  :: Docycle_method
        integer, optional,                 intent(out) :: STAT
        type(t_mgpu_transport_opts) :: o
        type(t_mgpu_geometry)       :: g
        type(t_mgpu_size3d)         :: s, w
        type(c_ptr)                 :: h
        integer                     :: rc, n
        o%ntracers = NTracers; o%precision = Precision; o%docycle_method = Docycle_method
        o%kernel = 0; o%tracers_per_launch = 0
        o%explicit_tracers_per_lane = 0; o%column_tracers_per_lane = 0; o%refinements = 0
        g = geometry_of(DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ)
        s = to_size3d(Size); w = to_size3d(WorkSize)
        TID = 0
        rc = mgpu_transport_create(s, w, o, g, h)
        if (rc == MGPU_SUCCESS) then
            do n = 1, MAX_INSTANCES
                if (.not. c_associated(THandles(n))) then
                    THandles(n) = h
                    TID = n
                    exit
                endif
            enddo
            if (TID == 0) then
                rc = mgpu_transport_destroy(h)
                rc = -99
            endif
        endif
        if (present(STAT)) STAT = rc
    end subroutine TransportStartGPU

    logical function bad_tid(TID, STAT)
        integer,           intent(in)  :: TID
        integer, optional, intent(out) :: STAT
        bad_tid = TID < 1 .or. TID > MAX_INSTANCES
        if (.not. bad_tid) bad_tid = .not. c_associated(THandles(TID))
        if (bad_tid .and. present(STAT)) STAT = MGPU_OFF_ERR
    end function bad_tid

    !> The geometry of the following steps (as SetGeometryGPU).
    subroutine TransportSetGeometryGPU(TID, DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ,   &
                                       DWZ, STAT)
        integer,                           intent(in)  :: TID
        real(8), dimension(:,:),   target, intent(in)  :: DUX, DVY, DZX, DZY
        integer, dimension(:,:),   target, intent(in)  :: KFloorZ
        real(8), dimension(:,:,:), contiguous, target, intent(in)  :: AreaU, AreaV, DZZ, DWZ
        integer, optional,                 intent(out) :: STAT
        integer :: rc
        if (bad_tid(TID, STAT)) return
        rc = mgpu_transport_set_geometry(THandles(TID),                                      &
                 geometry_of(DUX, DVY, DZX, DZY, KFloorZ, AreaU, AreaV, DZZ, DWZ))
        if (present(STAT)) STAT = rc
    end subroutine TransportSetGeometryGPU

    !> The hydrodynamic state and the transport options of the following
    !> steps: AdvectionDiffusion's arguments (ModuleAdvectionDiffusion.F90:
    !> 1108-1198) that every property of a step shares, with their names.
    subroutine TransportSetHydroGPU(TID, schmidt_H, SchmidtCoef_V, SchmidtBackground_V,      &
                                    AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV,  &
                                    Upwind2H, Upwind2V, VolumeRelMax, DTProp, ImpExp_AdvV,   &
                                    ImpExp_DifV, ImpExp_AdvXX, ImpExp_AdvYY, ImpExp_DifH,    &
                                    NullDif, Wflux_X, Wflux_Y, Wflux_Z, VolumeZOld, VolumeZ, &
                                    OpenPoints3D, LandPoints3D, ComputeFacesU3D,             &
                                    ComputeFacesV3D, ComputeFacesW3D, Visc_H, Diff_V,        &
                                    WaterPoints3D, SmallDepths, NoDifFlux, NoFluxU, NoFluxV, &
                                    NoFluxW, Optimize, STAT)
        integer,                           intent(in)    :: TID
        real(8),                           intent(in)    :: schmidt_H, SchmidtCoef_V
        real(8),                           intent(in)    :: SchmidtBackground_V
        integer,                           intent(in)    :: AdvMethodH, TVDLimitationH
        integer,                           intent(in)    :: AdvMethodV, TVDLimitationV
        logical,                           intent(in)    :: Upwind2H, Upwind2V
        real(8),                           intent(in)    :: VolumeRelMax, DTProp
        real(8),                           intent(in)    :: ImpExp_AdvV, ImpExp_DifV
        real(8),                           intent(in)    :: ImpExp_AdvXX, ImpExp_AdvYY
        real(8),                           intent(in)    :: ImpExp_DifH
        logical,                           intent(in)    :: NullDif
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: Wflux_X, Wflux_Y, Wflux_Z
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: VolumeZOld, VolumeZ
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: OpenPoints3D, LandPoints3D
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: ComputeFacesU3D, ComputeFacesV3D
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: ComputeFacesW3D
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: Visc_H, Diff_V
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: WaterPoints3D
        logical, dimension(:,:), optional, intent(in)    :: SmallDepths
        logical, optional,                 intent(in)    :: NoDifFlux
        integer, dimension(:,:,:), optional, contiguous, target, intent(in) :: NoFluxU, NoFluxV
        integer, dimension(:,:,:), optional, contiguous, target, intent(in) :: NoFluxW
        logical, optional,                 intent(in)    :: Optimize
        integer, optional,                 intent(out)   :: STAT
        type(t_mgpu_ad_args) :: a
        integer              :: rc
        integer, allocatable, target :: small(:,:)
        if (bad_tid(TID, STAT)) return
        a%schmidt_H = schmidt_H
        a%SchmidtCoef_V = SchmidtCoef_V
        a%SchmidtBackground_V = SchmidtBackground_V
        a%AdvMethodH = AdvMethodH; a%TVDLimitationH = TVDLimitationH
        a%AdvMethodV = AdvMethodV; a%TVDLimitationV = TVDLimitationV
        a%Upwind2H = b2i(Upwind2H); a%Upwind2V = b2i(Upwind2V)
        a%VolumeRelMax = VolumeRelMax
        a%AdvectionNudging = 0; a%AdvectionNudgingCells = 0
        a%DTProp = DTProp
        a%ImpExp_AdvV = ImpExp_AdvV; a%ImpExp_DifV = ImpExp_DifV
        a%ImpExp_AdvXX = ImpExp_AdvXX; a%ImpExp_AdvYY = ImpExp_AdvYY
        a%ImpExp_DifH = ImpExp_DifH
        a%NullDif = b2i(NullDif)
        a%Wflux_X = c_loc(Wflux_X); a%Wflux_Y = c_loc(Wflux_Y); a%Wflux_Z = c_loc(Wflux_Z)
        a%VolumeZOld = c_loc(VolumeZOld); a%VolumeZ = c_loc(VolumeZ)
        a%OpenPoints3D = c_loc(OpenPoints3D); a%LandPoints3D = c_loc(LandPoints3D)
        a%ComputeFacesU3D = c_loc(ComputeFacesU3D)
        a%ComputeFacesV3D = c_loc(ComputeFacesV3D)
        a%ComputeFacesW3D = c_loc(ComputeFacesW3D)
        a%Visc_H = c_loc(Visc_H); a%Diff_V = c_loc(Diff_V)
        a%WaterPoints3D = c_loc(WaterPoints3D)
        a%Optimize = 0; if (present(Optimize)) a%Optimize = b2i(Optimize)
        a%FirstProperty = 1
        a%CellFluxes = 0; a%NoAdvFlux = 0
        a%ReferenceProp = c_null_ptr; a%BoundaryCondition = 0; a%DecayTime = -9.9e15_8
        a%BoundaryPoints2D = c_null_ptr; a%PROPOld = c_null_ptr
        a%NoDifFlux = 0
        if (present(NoDifFlux)) a%NoDifFlux = b2i(NoDifFlux)
        a%NoFluxU = c_null_ptr; a%NoFluxV = c_null_ptr; a%NoFluxW = c_null_ptr
        if (present(NoFluxU)) a%NoFluxU = c_loc(NoFluxU)
        if (present(NoFluxV)) a%NoFluxV = c_loc(NoFluxV)
        if (present(NoFluxW)) a%NoFluxW = c_loc(NoFluxW)
        a%SmallDepths = c_null_ptr
        if (present(SmallDepths)) then
            allocate(small(size(SmallDepths, 1), size(SmallDepths, 2)))
            small = merge(1, 0, SmallDepths)
            a%SmallDepths = c_loc(small)
        endif
        a%device_pointers = 0
        rc = mgpu_transport_set_hydro(THandles(TID), a)
        if (present(STAT)) STAT = rc
    end subroutine TransportSetHydroGPU

    !> Property Index (1..NTracers) takes the concentration PROP (over Size).
    subroutine TransportSetPropertyGPU(TID, Index, PROP, STAT)
        integer,                           intent(in)  :: TID, Index
        real(8), dimension(:,:,:), contiguous, target, intent(in) :: PROP
        integer, optional,                 intent(out) :: STAT
        integer :: rc
        if (bad_tid(TID, STAT)) return
        rc = mgpu_transport_set_tracer(THandles(TID), int(Index - 1, c_int32_t), c_loc(PROP), 0)
        if (present(STAT)) STAT = rc
    end subroutine TransportSetPropertyGPU

    !> The concentration of property Index (1..NTracers) into PROP.
    subroutine TransportGetPropertyGPU(TID, Index, PROP, STAT)
        integer,                           intent(in)    :: TID, Index
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: PROP
        integer, optional,                 intent(out)   :: STAT
        integer :: rc
        if (bad_tid(TID, STAT)) return
        rc = mgpu_transport_get_tracer(THandles(TID), int(Index - 1, c_int32_t), c_loc(PROP), 0)
        if (present(STAT)) STAT = rc
    end subroutine TransportGetPropertyGPU

    !> NSteps transport steps of every property (captured as a hipGraph with
    !> UseGraph); waits for them and returns what the reference would have
    !> stopped on (MGPU_ERR_OPTION, MGPU_ERR_PIVOT).
    subroutine TransportRunGPU(TID, NSteps, UseGraph, STAT)
        integer,           intent(in)  :: TID, NSteps
        logical, optional, intent(in)  :: UseGraph
        integer, optional, intent(out) :: STAT
        integer :: rc
        logical :: g
        if (bad_tid(TID, STAT)) return
        g = .true.; if (present(UseGraph)) g = UseGraph
        rc = mgpu_transport_run(THandles(TID), int(NSteps, c_int32_t), b2i(g))
        if (rc == MGPU_SUCCESS) rc = mgpu_transport_status(THandles(TID))
        if (present(STAT)) STAT = rc
    end subroutine TransportRunGPU

    subroutine TransportKillGPU(TID, STAT)
        integer,           intent(inout) :: TID
        integer, optional, intent(out)   :: STAT
        integer :: rc
        if (bad_tid(TID, STAT)) return
        rc = mgpu_transport_destroy(THandles(TID))
        THandles(TID) = c_null_ptr
        TID = 0
        if (present(STAT)) STAT = rc
    end subroutine TransportKillGPU

    subroutine KillAdvectionDiffusionGPU(ID, STAT)
        integer,           intent(inout) :: ID
        integer, optional, intent(out)   :: STAT
        integer :: rc
        rc = MGPU_OFF_ERR
        if (ID >= 1 .and. ID <= MAX_INSTANCES) then
            if (c_associated(Handles(ID))) then
                rc = mgpu_ad_destroy(Handles(ID))
                Handles(ID) = c_null_ptr
                ID = 0
            endif
        endif
        if (present(STAT)) STAT = rc
    end subroutine KillAdvectionDiffusionGPU

    !> THOMASZ_NewType2 (ModuleFunctions.F90:4026-4123) on the caller's
    !> D/E/F/TI arrays (Size-shaped), writing the water columns of RES.
    subroutine THOMASZ_GPU(Size, WorkSize, D, E, F, TI, WaterPoints, RES, STAT)
        integer, dimension(6),             intent(in)    :: Size, WorkSize
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: D, E, F, TI
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: WaterPoints
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: RES
        integer, optional,                 intent(out)   :: STAT
        integer :: rc
        rc = mgpu_thomasz(to_size3d(Size), to_size3d(WorkSize), c_loc(D), c_loc(E), c_loc(F), &
                          c_loc(TI), c_loc(WaterPoints), c_loc(RES), 0_c_int32_t, c_null_ptr)
        if (present(STAT)) STAT = rc
    end subroutine THOMASZ_GPU

    !> THOMAS_3D (ModuleFunctions.F90:3667-3870, the T_THOMAS form) with the
    !> reference's bounds and direction arguments on the caller's D/E/F/TI
    !> arrays (Size-shaped), writing the lines' rows JImin..JImax+1 of RES.
    !> STAT = MGPU_ERR_PIVOT where the reference stops on a zero pivot.
    subroutine THOMAS_3D_GPU(Size, IJmin, IJmax, JImin, JImax, Kmin, Kmax, di, dj,   &
                             D, E, F, TI, RES, STAT)
        integer, dimension(6),             intent(in)    :: Size
        integer,                           intent(in)    :: IJmin, IJmax, JImin, JImax
        integer,                           intent(in)    :: Kmin, Kmax, di, dj
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: D, E, F, TI
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: RES
        integer, optional,                 intent(out)   :: STAT
        integer :: rc
        rc = mgpu_thomas3d(to_size3d(Size), int(IJmin, c_int32_t), int(IJmax, c_int32_t),  &
                           int(JImin, c_int32_t), int(JImax, c_int32_t),                   &
                           int(Kmin, c_int32_t), int(Kmax, c_int32_t),                     &
                           int(di, c_int32_t), int(dj, c_int32_t), c_loc(D), c_loc(E),     &
                           c_loc(F), c_loc(TI), c_loc(RES), 0_c_int32_t, c_null_ptr)
        if (present(STAT)) STAT = rc
    end subroutine THOMAS_3D_GPU

    !> The continuity step of ModuleHydrodynamic (WaterFlux_Z_VarGrid with
    !> Volume_Z_New and Volume_Z_Old, else WaterFlux_Z_FixGrid;
    !> ModuleHydrodynamic.F90:43270-43343): WaterFlux_Z of every water column,
    !> level by level upwards, from the horizontal fluxes (Size-shaped arrays).
    subroutine WaterFluxZGPU(Size, WorkSize, WaterFlux_X, WaterFlux_Y, WaterFlux_Z,       &
                             ComputeFaces3D_U, ComputeFaces3D_V, WaterPoints3D, DT,        &
                             Volume_Z_New, Volume_Z_Old, STAT)
        integer, dimension(6),             intent(in)    :: Size, WorkSize
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: WaterFlux_X, WaterFlux_Y
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: WaterFlux_Z
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: ComputeFaces3D_U
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: ComputeFaces3D_V
        integer, dimension(:,:,:), contiguous, target, intent(in)    :: WaterPoints3D
        real(8),                           intent(in)    :: DT
        real(8), dimension(:,:,:), optional, contiguous, target, intent(in) :: Volume_Z_New
        real(8), dimension(:,:,:), optional, contiguous, target, intent(in) :: Volume_Z_Old
        integer, optional,                 intent(out)   :: STAT
        type(c_ptr) :: vn, vo
        integer :: rc
        vn = c_null_ptr
        vo = c_null_ptr
        if (present(Volume_Z_New)) vn = c_loc(Volume_Z_New)
        if (present(Volume_Z_Old)) vo = c_loc(Volume_Z_Old)
        rc = mgpu_waterflux_z(to_size3d(Size), to_size3d(WorkSize), vn, vo, c_loc(WaterFlux_X), &
                              c_loc(WaterFlux_Y), c_loc(WaterFlux_Z), c_loc(ComputeFaces3D_U),  &
                              c_loc(ComputeFaces3D_V), c_loc(WaterPoints3D), real(DT, c_double), &
                              merge(1_c_int32_t, 0_c_int32_t, present(Volume_Z_New)),           &
                              0_c_int32_t)
        if (present(STAT)) STAT = rc
    end subroutine WaterFluxZGPU

    !> ComputeVolumes of ModuleGeometry (ModuleGeometry.F90:4108-4231):
    !> VolumeZ = DWZ DUX DVY and the face volumes; VolumeW only when given.
    subroutine ComputeVolumesGPU(Size, WorkSize, DWZ, DUX, DVY, VolumeZ, VolumeU, VolumeV,   &
                                 VolumeW, STAT)
        integer, dimension(6),             intent(in)    :: Size, WorkSize
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: DWZ
        real(8), dimension(:,:),   contiguous, target, intent(in)    :: DUX, DVY
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: VolumeZ, VolumeU, VolumeV
        real(8), dimension(:,:,:), optional, contiguous, target, intent(inout) :: VolumeW
        integer, optional,                 intent(out)   :: STAT
        type(c_ptr) :: vw
        integer :: rc
        vw = c_null_ptr
        if (present(VolumeW)) vw = c_loc(VolumeW)
        rc = mgpu_compute_volumes(to_size3d(Size), to_size3d(WorkSize), c_loc(DWZ), c_loc(DUX), &
                                  c_loc(DVY), c_loc(VolumeZ), c_loc(VolumeU), c_loc(VolumeV),    &
                                  vw, 0_c_int32_t)
        if (present(STAT)) STAT = rc
    end subroutine ComputeVolumesGPU

    !> ComputeAreas of ModuleGeometry (ModuleGeometry.F90:4238-4327): AreaU
    !> and AreaV; with SZZ and the impermeability coefficients (indexed by k
    !> over Size) the reduced areas.
    subroutine ComputeAreasGPU(Size, WorkSize, DUZ, DVZ, DXX, DYY, AreaU, AreaV, SZZ,        &
                               Coef_U, CoefX_U, Coef_V, CoefX_V, STAT)
        integer, dimension(6),             intent(in)    :: Size, WorkSize
        real(8), dimension(:,:,:), contiguous, target, intent(in)    :: DUZ, DVZ
        real(8), dimension(:,:),   contiguous, target, intent(in)    :: DXX, DYY
        real(8), dimension(:,:,:), contiguous, target, intent(inout) :: AreaU, AreaV
        real(8), dimension(:,:,:), optional, contiguous, target, intent(in) :: SZZ
        real(8), dimension(:), optional, contiguous, target, intent(in) :: Coef_U, CoefX_U
        real(8), dimension(:), optional, contiguous, target, intent(in) :: Coef_V, CoefX_V
        integer, optional,                 intent(out)   :: STAT
        type(c_ptr) :: p(5)
        integer :: rc
        p = c_null_ptr
        if (present(SZZ)) then
            if (.not. (present(Coef_U) .and. present(CoefX_U) .and. present(Coef_V) .and.    &
                       present(CoefX_V))) then
                if (present(STAT)) STAT = MGPU_ERR_BADARG
                return
            endif
            p(1) = c_loc(SZZ)
            p(2) = c_loc(Coef_U); p(3) = c_loc(CoefX_U)
            p(4) = c_loc(Coef_V); p(5) = c_loc(CoefX_V)
        endif
        rc = mgpu_compute_areas(to_size3d(Size), to_size3d(WorkSize), c_loc(DUZ), c_loc(DVZ), &
                                c_loc(DXX), c_loc(DYY), c_loc(AreaU), c_loc(AreaV), p(1), p(2), &
                                p(3), p(4), p(5), 0_c_int32_t)
        if (present(STAT)) STAT = rc
    end subroutine ComputeAreasGPU

end module ModuleAdvectionDiffusionGPU
