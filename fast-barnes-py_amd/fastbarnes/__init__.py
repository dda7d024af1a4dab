"""
fastbarnes -- fast Barnes interpolation on AMD Instinct MI355X (gfx950).

Drop-in for MeteoSwiss fast-barnes-py's ``fastbarnes.interpolation``; the
compute runs in HIP kernels of libfastbarnes.so (see include/fastbarnes.h).
"""
__version__ = '2.0.0+mi355x.1'
