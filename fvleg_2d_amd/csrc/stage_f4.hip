// Fused stage kernel instances for flux id 4.
#define CATO_STAGE_FLUX 4
#define CATO_STAGE_FN stage_launcher_f4
#include "stage_instances.inc"
