"""The Fortran drop-in keeps the reference's interface (CPU: compile only).

ModuleWaterProperties calls AdvectionDiffusion with a fixed text
(Software/MOHIDWater/ModuleWaterProperties.F90:14776-14822): 31 positional
arguments, then CellFluxes, WaterPoints3D, ReferenceProp, BoundaryCondition,
DecayTime, NumericStability, PropOld, SmallDepths, NoAdvFlux, NoDifFlux,
NoFluxU/V/W, Optimize, FirstProperty and STAT by keyword, from derived-type
components (ModuleWaterProperties' T_Property, T_External, T_SmallDepths,
T_NoFlux).  The switch to the GPU renames the call and adds the one argument
the reference fetches inside the call (BoundaryPoints2D, GetBoundaries,
AD:1373).  This test compiles that call, with stand-in derived types of the
same component names and kinds, against ModuleAdvectionDiffusionGPU, so an
argument out of the reference's order or a keyword the module lacks fails
here.  Nothing is linked or run.
"""
import os
import subprocess

import pytest

from mohid_amd import build as B

CALL = """
module wp_standins
    implicit none
    type T_AdvDifParameters
        real(8) :: SchmidtNumberH, SchmidtCoefV, SchmidtBackgroundV, VolumeRelMax
        real(8) :: AdvectionV_imp_exp, DiffusionV_imp_exp, DiffusionH_imp_exp, DecayTime
        integer :: AdvMethodH, TVDLimitationH, AdvMethodV, TVDLimitationV
        integer :: AdvectionNudgingCells, BoundaryCondition
        logical :: Upwind2H, Upwind2V, AdvectionNudging, NullDif, NumericStability
        real(8), pointer, dimension(:,:,:) :: PropOld => null()
    end type
    type T_Evolution
        type(T_AdvDifParameters) :: Advec_Difus_Parameters
        real(8) :: DTInterval
        logical :: NoAdvFluxCells, NoDifFluxCells, Discharges
    end type
    type T_Assimilation
        real(8), pointer, dimension(:,:,:) :: Field => null()
    end type
    type T_Property
        real(8), pointer, dimension(:,:,:) :: Concentration => null()
        type(T_Evolution)    :: evolution
        type(T_Assimilation) :: Assimilation
    end type
    type T_External
        real(8), pointer, dimension(:,:,:) :: WaterFluxX, WaterFluxY, WaterFluxZ
        real(8), pointer, dimension(:,:,:) :: VolumeZOld, VolumeZ, Visc_H, Diff_V
        integer, pointer, dimension(:,:,:) :: LandPoints3D, WaterPoints3D
        integer, pointer, dimension(:,:,:) :: ComputeFacesU3D, ComputeFacesV3D, ComputeFacesW3D
        integer, pointer, dimension(:,:)   :: BoundaryPoints2D
    end type
    type T_SmallDepths
        logical, pointer, dimension(:,:) :: ON => null()
    end type
    type T_NoFlux
        integer, pointer, dimension(:,:,:) :: U, V, W
    end type
    type T_WaterProperties
        integer              :: ObjAdvectionDiffusion
        type(T_External)     :: ExternalVar
        type(T_SmallDepths)  :: SmallDepths
        type(T_NoFlux)       :: NoFlux
    end type
end module wp_standins

subroutine Advection_Diffusion_Processes_GPU(Me, Property, OpenPoints3D, ImpExp_AdvXX,     &
                                             ImpExp_AdvYY, PropertyComputeBoxTimeSerie,   &
                                             OptimizeFlag, FirstWaterProperty)
    use ModuleAdvectionDiffusionGPU
    use wp_standins
    implicit none
    type(T_WaterProperties)            :: Me
    type(T_Property)                   :: Property
    integer, pointer, dimension(:,:,:) :: OpenPoints3D
    real(8)                            :: ImpExp_AdvXX, ImpExp_AdvYY
    logical                            :: PropertyComputeBoxTimeSerie
    logical                            :: OptimizeFlag, FirstWaterProperty
    integer                            :: STAT_CALL

    call AdvectionDiffusionGPU(Me%ObjAdvectionDiffusion,                                   &
            Property%Concentration,                                                        &
            Property%evolution%Advec_Difus_Parameters%SchmidtNumberH,                      &
            Property%evolution%Advec_Difus_Parameters%SchmidtCoefV,                        &
            Property%evolution%Advec_Difus_Parameters%SchmidtBackgroundV,                  &
            Property%evolution%Advec_Difus_Parameters%AdvMethodH,                          &
            Property%evolution%Advec_Difus_Parameters%TVDLimitationH,                      &
            Property%evolution%Advec_Difus_Parameters%AdvMethodV,                          &
            Property%evolution%Advec_Difus_Parameters%TVDLimitationV,                      &
            Property%evolution%Advec_Difus_Parameters%Upwind2H,                            &
            Property%evolution%Advec_Difus_Parameters%Upwind2V,                            &
            Property%evolution%Advec_Difus_Parameters%VolumeRelMax,                        &
            Property%evolution%Advec_Difus_Parameters%AdvectionNudging,                    &
            Property%evolution%Advec_Difus_Parameters%AdvectionNudgingCells,               &
            Property%evolution%DTInterval,                                                 &
            Property%evolution%Advec_Difus_Parameters%AdvectionV_imp_exp,                  &
            Property%evolution%Advec_Difus_Parameters%DiffusionV_imp_exp,                  &
            ImpExp_AdvXX, ImpExp_AdvYY,                                                    &
            Property%evolution%Advec_Difus_Parameters%DiffusionH_imp_exp,                  &
            Property%evolution%Advec_Difus_Parameters%NullDif,                             &
            Me%ExternalVar%WaterFluxX,                                                     &
            Me%ExternalVar%WaterFluxY,                                                     &
            Me%ExternalVar%WaterFluxZ,                                                     &
            Me%ExternalVar%VolumeZOld, Me%ExternalVar%VolumeZ,                             &
            OpenPoints3D,                                                                  &
            Me%ExternalVar%LandPoints3D,                                                   &
            Me%ExternalVar%ComputeFacesU3D,                                                &
            Me%ExternalVar%ComputeFacesV3D,                                                &
            Me%ExternalVar%ComputeFacesW3D,                                                &
            Me%ExternalVar%Visc_H,                                                         &
            Me%ExternalVar%Diff_V,                                                         &
            CellFluxes        = PropertyComputeBoxTimeSerie,                               &
            WaterPoints3D     = Me%ExternalVar%WaterPoints3D,                              &
            ReferenceProp     = Property%Assimilation%Field,                                &
            BoundaryCondition = Property%evolution%Advec_Difus_Parameters%BoundaryCondition, &
            DecayTime         = Property%evolution%Advec_Difus_Parameters%DecayTime,       &
            NumericStability  = Property%evolution%Advec_Difus_Parameters%NumericStability, &
            PropOld           = Property%evolution%Advec_Difus_Parameters%PropOld,         &
            SmallDepths       = Me%SmallDepths%ON,                                         &
            NoAdvFlux         = Property%evolution%NoAdvFluxCells,                         &
            NoDifFlux         = Property%evolution%NoDifFluxCells,                         &
            NoFluxU           = Me%NoFlux%U,                                               &
            NoFluxV           = Me%NoFlux%V,                                               &
            NoFluxW           = Me%NoFlux%W,                                               &
            Optimize          = OptimizeFlag,                                              &
            FirstProperty     = FirstWaterProperty,                                        &
            STAT              = STAT_CALL,                                                 &
            BoundaryPoints2D  = Me%ExternalVar%BoundaryPoints2D)
end subroutine Advection_Diffusion_Processes_GPU
"""

# the reference's positional prefix with one argument swapped (CellFluxes and
# WaterPoints3D in the order the previous version of the module had them):
# must not compile, so the test above is known to see argument order
SWAPPED = CALL.replace(
    "            Me%ExternalVar%Diff_V,                                                         &\n"
    "            CellFluxes        = PropertyComputeBoxTimeSerie,                               &\n"
    "            WaterPoints3D     = Me%ExternalVar%WaterPoints3D,                              &\n",
    "            Me%ExternalVar%Diff_V,                                                         &\n"
    "            Me%ExternalVar%WaterPoints3D,                                                  &\n"
    "            CellFluxes        = PropertyComputeBoxTimeSerie,                               &\n")


def _compile(tmp_path, text):
    mod = os.path.join(B.FDIR, "ModuleAdvectionDiffusionGPU.F90")
    src = tmp_path / "wp_call.F90"
    src.write_text(text)
    r = subprocess.run([B.FC, *B.FFLAGS, "-c", mod, "-o", str(tmp_path / "m.o"), "-J",
                        str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return subprocess.run([B.FC, *B.FFLAGS, "-c", str(src), "-o", str(tmp_path / "c.o"), "-J",
                           str(tmp_path)], capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module")
def fc():
    if not os.path.exists(B.FC):
        pytest.skip(f"no Fortran compiler at {B.FC}")
    return B.FC


def test_waterproperties_call_compiles(fc, tmp_path):
    r = _compile(tmp_path, CALL)
    assert r.returncode == 0, r.stdout + r.stderr


def test_argument_order_is_checked(fc, tmp_path):
    assert SWAPPED != CALL
    r = _compile(tmp_path, SWAPPED)
    assert r.returncode != 0
