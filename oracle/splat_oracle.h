/*
 * splat_oracle.h -- CPU oracle of the surfel splat prediction and fill-in
 * (IndexMap::combinedPredict, FillIn).  TEST INFRASTRUCTURE ONLY; see
 * splat_oracle.c for what it restates and the choices it makes.
 */
#ifndef SPLAT_ORACLE_H
#define SPLAT_ORACLE_H

#include <stdint.h>

#define ORC_SPLAT_FLOATS 12
#define ORC_SPLAT_MAX_POINT_SIZE 256

typedef struct {
    int width, height;
    float cx, cy, fx, fy;
    float depth_cutoff, conf_threshold;
    int time, max_time, time_delta;
} orc_splat_params;

/* model: n surfels of 12 floats {x, y, z, conf, colour, -, init_time,
 * last_time, nx, ny, nz, radius}; outputs width*height RGBA8, float4,
 * float4, uint16. */
void orc_splat_predict(const orc_splat_params *P, const float pose[16], const float *model, int n,
                       uint8_t *image, float *vertex, float *normal, uint16_t *time_out);

void orc_splat_fill_in(const orc_splat_params *P, const uint8_t *image, const float *vertex, const float *normal,
                       const uint16_t *depth, const uint8_t *rgb, int rgb_channels, int passthrough_geometry,
                       int passthrough_image, uint8_t *fimage, float *fvertex, float *fnormal);

#endif
