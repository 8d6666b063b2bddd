"""cubokit_amd — the MI355X (gfx950) `.hip` backend of CuboKit's first-processing hot path.

The product is the C-ABI library libcubo_hip.so (include/cubo_hip.h); this package builds it and mirrors
CuboKit's CuboVirtual API over it.
"""
from .origins import OriginEstimator, OriginMeasurement, ProbeEstimate, ProbeSize
from .virtual import (CancellationError, CenterOfMass, CenterOfMassMaps, ComputePath, CubeShape, CuboError,
                      CuboVirtualError, DetectorGeometry, DetectorImage, DetectorMask, FirstLook, HIPResidentCube, MaskStack,
                      ShiftedStack, StackBatch, center_of_mass_from_stack,
                      ResidentCubeError, ScanImage, ScanTiling, UnavailableReason, VirtualDiffraction, VirtualImaging,
                      device_count, device_info, first_look, hip_is_available, reduce_path,
                      DiffractionStatistic, OriginMap, PatternShift, ScanSelection, pattern_path,
                      HIPStreamedCube, NPYFile, NPYHeader, MaskShift, image_centered)

__all__ = ["CancellationError", "CenterOfMass", "CenterOfMassMaps", "ComputePath", "CubeShape", "CuboError",
           "CuboVirtualError", "DetectorGeometry", "DetectorImage", "DetectorMask", "FirstLook", "HIPResidentCube",
           "MaskStack", "ShiftedStack", "StackBatch", "center_of_mass_from_stack",
           "ResidentCubeError", "ScanImage", "ScanTiling", "UnavailableReason", "VirtualDiffraction",
           "VirtualImaging", "device_count", "device_info", "first_look", "hip_is_available", "reduce_path",
           "DiffractionStatistic", "OriginMap", "PatternShift", "ScanSelection", "pattern_path",
           "HIPStreamedCube", "NPYFile", "NPYHeader", "MaskShift", "image_centered",
           "OriginEstimator", "OriginMeasurement", "ProbeEstimate", "ProbeSize"]
