"""redrock_amd — redrock's redshift chi2 scan on AMD MI355X (gfx950).

Python host code mirroring redrock's chi2-scan API (zscan, rebin, igm, fitz,
zfind, targets, templates) on top of hand-written HIP kernels in
libredrock_hip.so, called through the C ABI declared in include/redrock_hip.h.
"""
__version__ = '0.1.0'
