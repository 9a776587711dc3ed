"""MI355X-native exact beam tracer (Overbeck, Ramamoorthi & Mark, EGSR 2007).

The GPU path is lib/libbtg.so (hand-written gfx950 HIP kernels behind the C ABI
in include/btg.h); the host input stage is lib/libbt_host.so (include/bth.h).
"""
