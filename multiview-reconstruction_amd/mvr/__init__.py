"""MI355X-native multiview Richardson-Lucy deconvolution (host side).

``mvr.hip`` binds the C ABI of ``libmvr_hip.so`` (HIP kernels for gfx950);
``mvr.decon`` mirrors the reference's MultiViewDeconvolution / DeconView /
ComputeBlockThread surface on top of it; ``mvr.synthetic`` generates seeded
inputs.  Importing this package loads nothing native.
"""
