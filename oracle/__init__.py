"""CPU oracle for the FujiShaderGPU raster hot path -- TEST INFRASTRUCTURE ONLY.

This package restates the reference's per-tile algorithm code on the CPU with
``cupy`` -> ``numpy`` and ``cupyx.scipy.ndimage`` -> ``scipy.ndimage``, keeping
the reference's float32 dtypes and operation order.  It is the checker that the
HIP path is compared against, and the CPU baseline that ``bench.py`` times
beside the GPU.  It is never imported by the product package
(``fujishadergpu_amd``); only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import it.

Modules
-------
ref    float32-faithful restatement of the reference block functions

Pinning (tests/test_oracle_pins.py checks each function against something
other than itself; the GPU tests then compare the HIP path with this oracle):

==========================  ==================================================
gradient / hillshade/slope  exact on planes with anisotropic steps and on
                            quadratics (edge_order=2); flat hillshade =
                            sin(alt); Lambert closed form on planes, both row
                            orientations; slope in every unit on planes
curvature                   reference tests/test_curvature_analytic.py values;
                            closed-form mean (bowl) and Gaussian (elliptic)
                            curvature at an apex
NaN Gaussian / box means    brute-force FP64 normalised convolutions with the
                            'nearest' and 'reflect' index rules
topousm                     z = x^2 -> -sum w r(r+1)/3; planes and constants
                            -> 0; downsample-factor table
decimate / zoom / sample    brute-force cell means, enclosed-void fill,
                            corner-aligned and pixel-centre bilinear loops;
                            integer overview block mean
openness                    reference tests/test_openness_yokoyama.py
                            orderings and flat = 1; cone-apex closed forms
                            (pi/2 - atan k), physical pixel distance
ambient occlusion           flat -> 1; FP64 ring average on a plane
stats / stretch             reference tests/test_topousm_fast_normalization
                            vectors; linear order statistics; stretch
quantize / NoData           reference tests/test_audit_p1_regressions.py and
                            tests/test_output_nodata_policy.py vectors;
                            round-half-even and clip rules
==========================  ==================================================

Parity unpinned (no pin exists without CuPy on an NVIDIA GPU, or GDAL):
the exact float32 bits of CuPy's gaussian_filter / uniform_filter / zoom
accumulation and of its float32 cell-sum reduction in _downsample_nan_aware
(restated with SciPy semantics: double accumulation, float32 per pass), the
precision of the 0-d promotion in compute_hillshade_block (restated in
float64, checked only against the closed form to 1e-6), and GDAL's
average-resampled overview read for a non-integer decimation
(overview_block_mean raises NotImplementedError there).  The GPU parity
tests state the tolerance that covers each.
"""
