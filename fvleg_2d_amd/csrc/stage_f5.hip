// Fused stage kernel instances for flux id 5.
#define CATO_STAGE_FLUX 5
#define CATO_STAGE_FN stage_launcher_f5
#include "stage_instances.inc"
