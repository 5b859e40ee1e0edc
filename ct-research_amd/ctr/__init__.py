"""ctr - MI355X (gfx950) CT projector pair, FDK/FBP and SIRT.

Host-side mirror of the CT-Research MEX operators over the C ABI of
libctr.so (include/ctr/ctr.h).  Every operator runs in HIP kernels; a
missing library raises instead of falling back to any CPU path.
"""
from ._lib import (CNN_LAYERS, CNN_NUM_PARAMS, Cnn, cnn_init_params, cnn_split_params)
from ._lib import (ACCUMULATE, AUTOTUNE, DEVICE_PTRS, FAN_EQUIANGULAR, FAN_FLAT, FILTER_RL_HALF, FILTER_SL_HALF,
                   REF_QUIRKS, ConeGeom, Context, CtrError, CurvedGeom, FanGeom, FdkErrGeom, ParGeom, default_context, lib, version, LIB_PATH,
                   EXPORTS)
from . import checkpoint, geometry, metrics, phantoms, uidata
from .mex import (FDK, Backprojection, Backprojection3D, BackprojectionCone_3D, BackprojectionFan2D,
                  BackprojectionFan_2D, FBPfan, FBPparallel, ProjectionCone_3D, ProjectionFan_2D,
                  ProjectionParallel_2D, SIRT3D, ImplementUI,
                  IterativeFDKerrCompute)
