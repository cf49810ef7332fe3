// Fused stage kernel instances for flux id 0.
#define CATO_STAGE_FLUX 0
#define CATO_STAGE_FN stage_launcher_f0
#include "stage_instances.inc"
