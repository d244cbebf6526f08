/* orkan.h -- umbrella header of the Orkan C API (reference include/orkan.h). */
#ifndef ORKAN_H
#define ORKAN_H
#include "q_types.h"
#include "state.h"
#include "gate.h"
#include "channel.h"
#include "meas.h"
#include "circ.h"
#include "sample.h"   /* extension: sample_matrix (specified, not implemented, by the reference) */
#endif /* ORKAN_H */
