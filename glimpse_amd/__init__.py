"""glimpse_amd: MI355X-native GLIMPSE2_phase compute core.

The product path is the HIP library libglimpse_hip.so behind the C-ABI in
include/glimpse_hip.h; this package only loads it (glimpse_amd.hip) and
provides the synthetic-data front-end (glimpse_amd.synth)."""
