"""metriko_amd: MI355X-native (gfx950) HIP implementation of metriko's field + IGM hot path."""
