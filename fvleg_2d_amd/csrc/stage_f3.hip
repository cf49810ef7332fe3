// Fused stage kernel instances for flux id 3.
#define CATO_STAGE_FLUX 3
#define CATO_STAGE_FN stage_launcher_f3
#include "stage_instances.inc"
