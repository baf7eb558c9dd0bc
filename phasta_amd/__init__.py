"""phasta_amd -- MI355X-native compressible assembly and block-GMRES.

The product is ``libphgpu.so`` (HIP, gfx950) behind the C-ABI in
``include/phgpu.h``; this package holds the Python-side mirror of the
reference's ``ElmGMRs``/``SolGMRs`` interface used by tests and the bench,
and the synthetic box-mesh builder.
"""
