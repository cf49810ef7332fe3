// Fused stage kernel instances for flux id 8.
#define CATO_STAGE_FLUX 8
#define CATO_STAGE_FN stage_launcher_f8
#include "stage_instances.inc"
