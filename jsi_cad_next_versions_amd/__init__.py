"""MI355X-native minimum-distance and surface-surface-intersection hot path
of JSI_CAD_NEXT (VERSION5 query API), backed by hand-written HIP kernels in
libjsi_mi355x.so."""
