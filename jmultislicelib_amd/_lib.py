"""ctypes binding of libjmultislice_hip.so (the MI355X multislice library).

The shared library is built in-tree by ``jmultislicelib_amd/csrc/Makefile``
(``__graft_entry__.build()``).  Loading fails loudly when it is missing: the
package has no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libjmultislice_hip.so")

_lib = None

c_int = C.c_int
c_float = C.c_float
c_double = C.c_double
c_void_p = C.c_void_p
c_char_p = C.c_char_p
c_longlong = C.c_longlong
c_uint = C.c_uint

# (name, restype, argtypes) of every exported entry point
_SIGS = [
    # flat C-ABI (include/JMultiSliceLib.h)
    ("GetJMSVersion", c_float, []),
    ("SetJMSDebugLevel", c_int, [c_int]),
    ("SetRng", None, [c_void_p]),
    ("GetRng", c_void_p, []),
    ("SeedRngEx", None, [c_int]),
    ("SetHighTension", c_float, [c_float]),
    ("GetHighTension", c_float, []),
    ("GetWaveLength", c_float, []),
    ("SetSupercellSize", None, [c_void_p]),
    ("SetSupercellSizeABC", None, [c_float, c_float, c_float]),
    ("SetSliceThickness", None, [c_int, c_float]),
    ("DiffractionDescan", None, [c_int]),
    ("SetPlasmonMC", None, [c_int, c_float, c_float, c_float, C.c_uint]),
    ("SetDiffractionDescanN", None, [c_int, c_int, c_int, c_int]),
    ("SetDiffractionDescan", None, [c_int, c_float, c_float, c_int]),
    ("SetDiffractionDescanMRad", None, [c_int, c_float, c_float, c_int]),
    ("SetGridSize", None, [c_int, c_int]),
    ("GetGridSize", None, [c_void_p, c_void_p]),
    ("GetDetetionSliceNum", c_int, []),
    ("GetDetNum", c_int, []),
    ("GetCPUNum", c_int, []),
    ("GetGPUNum", c_int, []),
    ("GetGPUName", c_int, [c_int, c_char_p]),
    ("GetCurrentGPU", c_int, []),
    ("SetCurrentGPU", c_int, [c_int]),
    ("GetGPUStats", c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("GetUnscrambleHash", c_int, [c_void_p]),
    ("GetImageAveragingWeight", c_float, [c_int, c_int]),
    ("GetWaveAveragingWeight", c_float, [c_int, c_int]),
    ("PhaseGratingSetup", c_int, [c_int, c_int, c_int, c_int, c_int, c_void_p]),
    ("ObjectSliceSetup", c_int, [c_int, c_void_p]),
    ("PropagatorSetup", c_int, [c_int, c_int, c_void_p]),
    ("DetectorSetup", c_int, [c_int, c_int, c_int, c_int, c_int]),
    ("InitCore", c_int, [c_int, c_int]),
    ("SetIncidentWave", c_int, [c_int, c_void_p, c_int]),
    ("SetIncidentWavePhaseplate", c_int, [c_int, c_void_p, c_int]),
    ("ZeroIncidentWavePhaseplate", c_int, [c_int]),
    ("SetSlicePhaseGratings", c_int, [c_int, c_int, c_int, c_void_p]),
    ("SetPropagatorData", c_int, [c_int, c_int, c_void_p]),
    ("SetDetectorData", c_int, [c_int, c_int, c_void_p, c_int, c_void_p]),
    ("SetGPUPgrLoading", c_int, [c_int]),
    ("ResetImageAveraging", c_int, [c_int, c_int]),
    ("ResetWaveAveraging", c_int, [c_int, c_int]),
    ("CalculateProbeWaveFourier", c_int, [c_void_p, c_void_p]),
    ("SetIncomingWaveGPU", c_int, [c_void_p, c_int]),
    ("SetIncomingWaveCPU", c_int, [c_void_p, c_int, c_int]),
    ("CalculatePropagator", c_int, [c_float, c_float, c_float, c_void_p, c_int]),
    ("CalculateRingDetector", c_int, [c_float, c_float, c_float, c_float, c_float, c_float, c_char_p,
                                      c_void_p, c_void_p, c_void_p, c_int]),
    ("OffsetIncomingWave", c_int, [c_int, c_float, c_float, c_float, c_int, c_int, c_int, c_int]),
    ("CPUMultislice", c_int, [c_int, c_int, c_float, c_int]),
    ("GPUMultislice", c_int, [c_int, c_int, c_float]),
    ("GetResult", c_int, [c_int, c_int, c_void_p, c_int]),
    ("GetAvgResult", c_int, [c_int, c_int, c_void_p, c_void_p, c_int]),
    ("ClearDetMem_h", c_int, [c_int]),
    ("ClearDetMem_d", c_int, []),
    ("GPUMultisliceScan", c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("GPUMultisliceScanVar", c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                     c_void_p]),
    ("GetRandomVariantSequence", c_int, [c_void_p]),
    ("Cleanup", c_int, []),
    ("FreeLibMem", None, []),
    # thin engine ABI (include/jms_hip.h)
    ("jms_hip_create", c_int, [c_void_p]),
    ("jms_hip_destroy", None, [c_void_p]),
    ("jms_hip_grid_supported", c_int, [c_int, c_int]),
    ("jms_hip_set_grid", c_int, [c_void_p, c_int, c_int, c_float, c_float, c_float]),
    ("jms_hip_setup_gratings", c_int, [c_void_p, c_int, c_int, c_int, c_void_p]),
    ("jms_hip_upload_gratings", c_int, [c_void_p, c_int, c_int, c_void_p]),
    ("jms_hip_gaussian_gratings", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double,
                                          c_double]),
    ("jms_hip_formfactor_gratings", c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p, c_float, c_int]),
    ("jms_hip_get_gratings", c_int, [c_void_p, c_int, c_int, c_void_p]),
    ("jms_hip_setup_propagators", c_int, [c_void_p, c_int, c_int, c_void_p]),
    ("jms_hip_upload_propagator", c_int, [c_void_p, c_int, c_void_p]),
    ("jms_hip_setup_objects", c_int, [c_void_p, c_int, c_void_p]),
    ("jms_hip_setup_detection", c_int, [c_void_p, c_int, c_void_p, c_int, c_int]),
    ("jms_hip_upload_detector", c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p]),
    ("jms_hip_set_probe", c_int, [c_void_p, c_void_p]),
    ("jms_hip_set_phaseplate", c_int, [c_void_p, c_void_p]),
    ("jms_hip_probe", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("jms_hip_stage_scan", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                   c_void_p]),
    ("jms_hip_run_staged", c_int, [c_void_p, c_int, c_int, c_int]),
    ("jms_hip_sync", c_int, [c_void_p]),
    ("jms_hip_run_batch_noalloc", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                          c_void_p, c_void_p]),
    ("jms_hip_group_create", c_int, [c_void_p, c_void_p, c_int]),
    ("jms_hip_group_run", c_int, [c_void_p]),
    ("jms_hip_group_last_error", c_char_p, [c_void_p]),
    ("jms_hip_group_destroy", None, [c_void_p]),
    ("jms_hip_get_scan_results", c_int, [c_void_p, c_void_p]),
    ("jms_hip_set_batch", c_int, [c_void_p, c_int]),
    ("jms_hip_set_graph", c_int, [c_void_p, c_int]),
    ("jms_hip_wave_offset", c_int, [c_void_p, c_float, c_float, c_float, c_int, c_int, c_int]),
    ("jms_hip_wave_set", c_int, [c_void_p, c_void_p]),
    ("jms_hip_wave_get", c_int, [c_void_p, c_void_p]),
    ("jms_hip_multislice", c_int, [c_void_p, c_int, c_void_p, c_float, c_int]),
    ("jms_hip_get_result", c_int, [c_void_p, c_int, c_void_p]),
    ("jms_hip_get_avg_result", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("jms_hip_get_averaging_weights", c_int, [c_void_p, c_void_p, c_void_p]),
    ("jms_hip_set_descan", c_int, [c_void_p, c_int, c_int]),
    ("jms_hip_reset_averaging", c_int, [c_void_p, c_int]),
    ("jms_hip_readout_det_avg", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("jms_hip_fft2d", c_int, [c_int, c_int, c_int, c_void_p, c_int]),
    ("jms_hip_set_timing", c_int, [c_void_p, c_int]),
    ("jms_hip_get_timing", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("jms_hip_last_error", c_char_p, [c_void_p]),
    # variant sequencing (include/jms_sequence.h)
    ("jms_variant_sequences", c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    # on-disk formats (include/jms_io.h)
    ("jms_ems_read_header", c_int, [c_char_p, c_void_p]),
    ("jms_ems_read_data", c_int, [c_char_p, c_void_p, c_void_p]),
    ("jms_ems_write", c_int, [c_char_p, c_void_p, c_void_p]),
    ("jms_stem_images", c_int, [c_int, c_int, c_int, c_void_p, c_float, c_void_p]),
    ("jms_bin_write_channel", c_int, [c_char_p, c_void_p, c_int, c_int, c_int, c_int]),
    # scattering factors and structure slicing (include/jms_elasca.h)
    ("jms_weko_params", c_int, [c_int, c_void_p, c_void_p]),
    ("jms_weko_f", c_double, [c_void_p, c_void_p, c_double]),
    ("jms_weko_ei", c_double, [c_double]),
    ("jms_weko_rih2", c_double, [c_double]),
    ("jms_weko_ri1", c_double, [c_double, c_double, c_double]),
    ("jms_weko_ri2", c_double, [c_double, c_double, c_double, c_double]),
    ("jms_weko_imag", c_double, [c_double, c_double, c_void_p, c_void_p]),
    ("jms_weko_form_factor", c_int, [c_double, c_double, c_int, c_double, c_int, c_int, c_void_p]),
    ("jms_radial_form_factors", c_int, [c_int, c_float, c_float, c_int, c_int, c_float, c_longlong, c_float,
                                        c_float, c_void_p]),
    ("jms_form_factor_sampling", c_int, [c_int, c_int, c_int, c_float, c_float, c_float, c_void_p, c_void_p,
                                         c_void_p]),
    ("jms_form_factors_2d", c_int, [c_int, c_int, c_float, c_float, c_int, c_int, c_longlong, c_float, c_void_p,
                                    c_void_p]),
    ("jms_zslice", c_int, [c_int, c_void_p, c_float, c_float, c_void_p, c_void_p]),
    ("jms_suggest_zslices", c_int, [c_int, c_void_p, c_void_p, c_float, c_float]),
    # specimen builder (include/jms_specimen.h)
    ("jms_perovskite_sites", c_int, [c_int, c_int, c_double, c_int, c_int, c_int, c_int, c_double, c_int,
                                     c_int, c_void_p, c_void_p]),
    ("jms_gaussian_phase_grating", c_int, [c_int, c_int, c_double, c_double, c_int, c_void_p, c_void_p,
                                           c_double, c_double, c_void_p]),
    # surrogate network (include/jms_nn.h)
    ("jms_nn_create", c_int, [c_void_p, c_int, c_uint]),
    ("jms_nn_set_kernels", c_int, [c_void_p, c_int]),
    ("jms_nn_destroy", None, [c_void_p]),
    ("jms_nn_num_params", c_int, []),
    ("jms_nn_get_params", c_int, [c_void_p, c_void_p]),
    ("jms_nn_set_params", c_int, [c_void_p, c_void_p]),
    ("jms_nn_forward", c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    ("jms_nn_gradients", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("jms_nn_data_create", c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    ("jms_nn_data_destroy", None, [c_void_p]),
    ("jms_nn_data_windows", c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_float, c_float, c_void_p]),
    ("jms_nn_data_get", c_int, [c_void_p, c_void_p]),
    ("jms_nn_train_epoch", c_int, [c_void_p, c_void_p, c_int, c_float, c_void_p, c_uint, c_void_p]),
    ("jms_nn_predict", c_int, [c_void_p, c_void_p, c_void_p]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib() -> C.CDLL:
    """Load (once) and return the library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray | None):
    """c_void_p of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


class CJProbeParams(C.Structure):
    """Probe parameters (include/JMultiSliceLib.h, fields of JProbeGen.h:84-122)."""
    _fields_ = [
        ("m_wl", c_float), ("m_alpha", c_float), ("m_alpha_rs", c_float),
        ("m_alpha_x0", c_float), ("m_alpha_y0", c_float), ("m_alpha_asym", c_float),
        ("m_alpha_adir", c_float), ("m_btx", c_float), ("m_bty", c_float),
        ("m_source_width", c_float), ("m_source_shape", c_int), ("m_fspread_width", c_float),
        ("m_fspread_kernel_width", c_float), ("m_fspread_kernel_samples", c_int),
        ("m_abrr_coeff", C.POINTER(c_float)), ("m_abrr_num", c_int),
    ]
