/* backend_hip_stub.c: backend_hip_create for encoder builds without HIP
 * (the convention of src/backend_metal_stub.c). */
#include <stdio.h>

#include "backend.h"

Backend *backend_hip_create(struct Mav1Encoder *e, char *err, size_t errlen) {
  (void)e;
  if (err && errlen) snprintf(err, errlen, "this build has no HIP backend");
  return NULL;
}
