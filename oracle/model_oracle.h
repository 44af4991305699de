/*
 * model_oracle.h -- CPU oracle of the surfel model's fusion (Model::initialise,
 * IndexMap::predictIndices, Model::fuse, Model::clean).  TEST INFRASTRUCTURE
 * ONLY; see model_oracle.c for what it restates and the choices it makes.
 * Surfels are 12 floats {x, y, z, confidence, colour, 0, init_time,
 * last_time, nx, ny, nz, radius}.
 */
#ifndef MODEL_ORACLE_H
#define MODEL_ORACLE_H

#include <stdint.h>

typedef struct orc_model orc_model;

orc_model *orc_model_create(int w, int h, float cx, float cy, float fx, float fy, int max_surfels);
void orc_model_destroy(orc_model *m);
int orc_model_count(const orc_model *m);
const float *orc_model_data(const orc_model *m);
int orc_model_fresh_count(const orc_model *m);
const float *orc_model_fresh(const orc_model *m);
int orc_model_set(orc_model *m, const float *model, int n);
void orc_model_index_map(const orc_model *m, uint32_t *index, float *vert_conf, float *color_time, float *norm_rad);
void orc_depth_metric(const uint16_t *d, int n, float maxD, float *out);
int orc_model_initialise(orc_model *m, const uint8_t *rgb, const uint16_t *raw, const uint16_t *filtered, float maxD,
                         int time, float maxDepth);
void orc_model_predict_indices(orc_model *m, const float pose[16], int time, float maxDepth, int timeDelta);
void orc_model_fuse(orc_model *m, const float pose[16], int time, const uint8_t *rgb, const uint16_t *raw,
                    const uint16_t *filtered, float maxD, float maxDepth, float weighting);
int orc_model_clean(orc_model *m, const float pose[16], int time, float conf_threshold, int time_delta);

#endif
