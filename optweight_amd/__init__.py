'''
optweight_amd: optweight's conjugate-gradient Wiener-filter path
(solvers / operators / preconditioners / sht / multigrid) on one MI355X,
with hand-written HIP kernels behind a C ABI (include/owmi.h,
libowmi.so).
'''
__version__ = '0.1.0'
