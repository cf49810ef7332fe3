// Fused stage kernel instances for flux id 1.
#define CATO_STAGE_FLUX 1
#define CATO_STAGE_FN stage_launcher_f1
#include "stage_instances.inc"
