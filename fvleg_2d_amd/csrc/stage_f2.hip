// Fused stage kernel instances for flux id 2.
#define CATO_STAGE_FLUX 2
#define CATO_STAGE_FN stage_launcher_f2
#include "stage_instances.inc"
