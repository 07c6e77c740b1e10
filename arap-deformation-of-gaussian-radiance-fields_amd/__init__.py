"""MI355X-native per-frame ARAP deform + render path (HIP kernels behind a C ABI).

The directory name is not a Python identifier, so callers load it under the
module name ``arapgs_amd`` (see ``load_package`` in bench.py / tests/conftest.py).
"""
