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
fp64   float64 / closed-form oracles used to pin ``ref`` (and exact modes)

Pinning (what checks each oracle function; see tests/test_oracle_*.py):

==========================  ==================================================
gradient / hillshade/slope  closed forms on planes; flat hillshade = sin(alt);
                            np.gradient on quadratics (exact); FP64 oracle
curvature                   quadratic surfaces (closed-form kappa everywhere);
                            reference tests/test_curvature_analytic.py values;
                            signed-step fixture
NaN Gaussian fill           brute-force FP64 9x9 double loop on small grids
box / topousm               z = x^2 closed form -sum w r(r+1)/3; planes -> 0;
                            brute-force box means on small grids
decimate / zoom / sample    brute-force loops on tiny grids (block means,
                            corner-aligned bilinear)
openness                    apex closed forms (flat, cone, pit); reference
                            tests/test_openness_yokoyama.py; brute-force FP64
ambient occlusion           flat -> 1; brute-force FP64 loop on small grids
percentile / stats          numpy.percentile(method='linear') in FP64;
                            reference tests/test_topousm_fast_normalization.py
quantize / NoData           reference tests/test_audit_p1_regressions.py
                            vectors; exhaustive-style sweeps vs the formula
==========================  ==================================================

Bit-level parity with CuPy's own float32 filter accumulation order
(gaussian_filter, uniform_filter, zoom, reductions) is not reproducible
without CuPy and an NVIDIA GPU; those calls are restated with SciPy's
semantics (double accumulation, float32 output per pass) and compared under a
stated tolerance.
"""
