/* Input side and quality measure of the command-line front end
 * (tools/mav1enc.c): the y4m (I420) reader of :67-77 and the PSNR of
 * :22-34, as library calls (include/mav1_hip.h, mav1hip_y4m_*,
 * mav1hip_psnr). Host code only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mav1_hip.h"

struct mav1hip_y4m {
  FILE *f;
  int w, h, cw, ch;
};

static void set_err(char *err, size_t errlen, const char *msg) {
  if (err && errlen) snprintf(err, errlen, "%s", msg);
}

mav1hip_y4m *mav1hip_y4m_open(const char *path, int32_t *width, int32_t *height, int32_t *fps_num, int32_t *fps_den,
                              char *err, size_t errlen) {
  FILE *f = path ? fopen(path, "rb") : NULL;
  if (!f) {
    set_err(err, errlen, "cannot open the input file");
    return NULL;
  }
  char line[512];
  if (!fgets(line, sizeof line, f) || strncmp(line, "YUV4MPEG2", 9)) {
    fclose(f);
    set_err(err, errlen, "not y4m");
    return NULL;
  }
  /* stream header tags: W, H, F (frame rate num:den), C (colour space,
   * only 4:2:0 accepted); the others are ignored */
  int w = 0, h = 0, fn = 30, fd = 1;
  char *save = NULL;
  for (char *t = strtok_r(line, " \n", &save); t; t = strtok_r(NULL, " \n", &save)) {
    if (t[0] == 'W')
      w = atoi(t + 1);
    else if (t[0] == 'H')
      h = atoi(t + 1);
    else if (t[0] == 'F')
      sscanf(t + 1, "%d:%d", &fn, &fd);
    else if (t[0] == 'C' && strncmp(t + 1, "420", 3)) {
      fclose(f);
      set_err(err, errlen, "only 4:2:0 input is supported");
      return NULL;
    }
  }
  if (w < 16 || h < 16 || w > 65536 || h > 65536) {
    fclose(f);
    set_err(err, errlen, "unsupported frame size");
    return NULL;
  }
  mav1hip_y4m *y = (mav1hip_y4m *)calloc(1, sizeof *y);
  if (!y) {
    fclose(f);
    set_err(err, errlen, "out of memory");
    return NULL;
  }
  y->f = f;
  y->w = w;
  y->h = h;
  y->cw = (w + 1) / 2;
  y->ch = (h + 1) / 2;
  if (width) *width = w;
  if (height) *height = h;
  if (fps_num) *fps_num = fn;
  if (fps_den) *fps_den = fd;
  return y;
}

int mav1hip_y4m_read(mav1hip_y4m *y, uint8_t *const planes[3], const int32_t strides[3]) {
  if (!y || !planes || !strides) return MAV1HIP_EINVAL;
  /* each frame: a "FRAME" line (parameters ignored), then Y, U and V */
  char line[256];
  if (!fgets(line, sizeof line, y->f)) return 1;
  if (strncmp(line, "FRAME", 5)) return MAV1HIP_EINVAL;
  for (int p = 0; p < 3; p++) {
    const int pw = p ? y->cw : y->w, ph = p ? y->ch : y->h;
    if (!planes[p] || strides[p] < pw) return MAV1HIP_EINVAL;
    for (int r = 0; r < ph; r++)
      if (fread(planes[p] + (size_t)r * strides[p], 1, (size_t)pw, y->f) != (size_t)pw) return MAV1HIP_EINVAL;
  }
  return 0;
}

void mav1hip_y4m_close(mav1hip_y4m *y) {
  if (!y) return;
  if (y->f) fclose(y->f);
  free(y);
}

double mav1hip_psnr(const uint8_t *a, const uint8_t *b, size_t n) {
  /* exact integer sum of squared differences */
  unsigned long long sse = 0;
  for (size_t i = 0; i < n; i++) {
    const int d = (int)a[i] - (int)b[i];
    sse += (unsigned long long)(d * d);
  }
  if (sse == 0) return 99.0;
  return 10.0 * log10(255.0 * 255.0 * (double)n / (double)sse);
}
