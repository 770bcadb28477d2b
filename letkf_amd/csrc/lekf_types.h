/* lekf_types.h -- device-side alias of the public job descriptor. */
#ifndef LEKF_TYPES_H
#define LEKF_TYPES_H
#include "lekf.h"
typedef lekf_job LekfJob;
#endif
