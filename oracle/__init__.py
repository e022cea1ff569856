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
ref       float32-faithful restatement of the reference block functions
geometry  tile grid, tile halo, valid-data box and stratified pre-pass windows
prepare   the prepare step (NoData inference, push-pull fill)
cnn       FP64 reference of the openness surrogate network
tiff      pure-Python TIFF / BigTIFF reader

Pinning (tests/test_oracle_pins.py checks each function against something
other than itself; the GPU tests then compare the HIP path with this oracle):

==========================  ==================================================
gradient / hillshade/slope  exact on planes with anisotropic steps and on
                            quadratics (edge_order=2); flat hillshade =
                            sin(alt); Lambert closed form on planes, both row
                            orientations; slope in every unit on planes
curvature                   reference tests/test_curvature_analytic.py values;
                            closed-form mean (bowl) and Gaussian (elliptic)
                            curvature at an apex; the signed row step (bowl
                            apex H = 0, saddle s = 0 on north-up rasters)
NaN Gaussian / box means    brute-force FP64 normalised convolutions with the
                            'nearest' and 'reflect' index rules
topousm                     z = x^2 -> -sum w r(r+1)/3; planes and constants
                            -> 0; downsample-factor table
decimate / zoom / sample    brute-force cell means, enclosed-void fill,
                            corner-aligned and pixel-centre bilinear loops;
                            integer overview block mean
overview (any ratio)        supersampled block means at rational ratios,
                            constants, mass conservation, float64 block mean
                            at integer ratios
large-radius field          constant -> 0, a ramp -> 0 inside (the window
                            offset and the pixel-centre map), the r / s
                            rounding table with its sigma = 1 branch
spatial smoothing           sigma = r / 2 from the second moment of u^2
geometry                    hand-written tile and pre-pass window lists for
                            the 16384^2, 32768^2 and ragged rasters
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
float64, checked only against the closed form to 1e-6), and GDAL's own
average resampling (overview_average restates it as the area-weighted mean
of the full-resolution pixels; GDAL may resample from an internal overview
level instead, and rounds in its own order).  The GPU parity tests state the
tolerance that covers each.
"""
