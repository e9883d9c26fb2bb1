/* gpu_matrix.h -- included by elegant under HAVE_GPU; every declaration of the
 * adapter is in gpu_base.h */
#include "gpu_base.h"
