// Fused stage kernel instances for flux id 6.
#define CATO_STAGE_FLUX 6
#define CATO_STAGE_FN stage_launcher_f6
#include "stage_instances.inc"
