"""MI355X-native backend of the SPRING memetic phase-retrieval engine.

The hot path (per-individual HIO/RAAR/ER iterations, shrink-wrap, fitness,
alignment and crossover of a population) runs as hand-written gfx950 HIP
kernels in lib/libspring_hip.so, reached through the C ABI of
include/spring_hip.h (binding: s_spi_spring_amd.hipcore).
"""
