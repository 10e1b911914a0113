"""apollo_amd: MI355X-native nodal-DG right-hand-side engine for Apollo's
`nodalDG2d` subsolver (Euler and Maxwell on unstructured triangles).

The compute path is the HIP library ``libapollo_dg.so`` behind the C-ABI in
``include/apollo_dg.h``; ``apollo_amd.engine`` is its ctypes binding.
"""
from . import engine  # noqa: F401
