// Fused stage kernel instances for flux id 7.
#define CATO_STAGE_FLUX 7
#define CATO_STAGE_FN stage_launcher_f7
#include "stage_instances.inc"
