// femrt_march_ng7.hip -- the element walk kernels for 7 geometry modes per
// direction (all field mode counts, learned/unlearned, counted/timed,
// cooperative/sequential scan).
#include "femrt_march.hpp"

FEMRT_MARCH_LAUNCH_DEF(7)
