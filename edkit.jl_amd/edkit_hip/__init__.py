"""edkit_hip -- MI355X (gfx950) engine behind EDKit.jl's matrix-free Operator
path, driven from Python through the C ABI of libedkit_hip.so.

The names mirror the Julia API: bases (TensorBasis, ProjectedBasis,
TranslationalBasis, TranslationParityBasis, ParityBasis, FlipBasis,
ParityFlipBasis, TranslationFlipBasis, basis), spin, operator,
trans_inv_operator, mul_ (mul!), H @ x / H * x, mul, timeevolve,
KrylovEvolutionCache, timeevolve_ (timeevolve!), productstate, schmidt,
ent_spec, ent_S, DoubleBasis, symmetrizer, basis_embedding, lindblad, densitymatrix, LiouvillianMap,
LindbladArnoldiCache, lindblad_timeevolve, lindblad_timeevolve_
(lindblad_timeevolve!), plus the new lanczos_ground.
"""
from ._lib import EdkError, lib  # noqa: F401
from .batch import MatvecBatch  # noqa: F401
from .schmidt import ent_S, ent_spec, entropy, schmidt  # noqa: F401
from .basis import (AbelianBasis, AbstractBasis, FlipBasis, ParityBasis,  # noqa: F401
                    ParityFlipBasis, ProjectedBasis, TensorBasis, TranslationalBasis,
                    TranslationFlipBasis, TranslationParityBasis, basis, change_, content, index, norm,
                    productstate)
from .device import Context, DeviceArray, default_context, set_default_context  # noqa: F401
from .linearmap import DoubleBasis, basis_embedding, symmetrizer  # noqa: F401
from .lindblad import (DensityMatrix, Lindblad, LindbladArnoldiCache, LiouvillianMap,  # noqa: F401
                       densitymatrix, lindblad, lindblad_timeevolve, lindblad_timeevolve_)
from .krylov import KrylovEvolutionCache, lanczos_ground, timeevolve, timeevolve_  # noqa: F401
from .matvec import (Array, addto_, apply, clear_sparse_cache_, device_operator, matvec_into, mul,  # noqa: F401
                     mul_, mul_multi, sparse, sparse_)
from .operator import Operator, find_base, operator, trans_inv_operator  # noqa: F401
from .serve import Server  # noqa: F401
from .spin import spin  # noqa: F401
