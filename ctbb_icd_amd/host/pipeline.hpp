// pipeline.hpp -- reference-named pipeline stages (see pipeline.cpp).
#pragma once

#include "recon_structs.hpp"

void load_raw(struct recon_params *rp, struct ct_data *data);
void initialize_recon_volume(const struct recon_params *rp, struct ct_data *data);
void generate_system_matrix(const struct recon_params *rp, struct ct_data *data);
void generate_system_matrix_ffs(const struct recon_params *rp, struct ct_data *data);
void load_system_matrix(const struct recon_params *rp, struct ct_data *data);
void icd_iteration(const struct recon_params *rp, struct ct_data *data);
void rotate_slices_fixed2rotating(const struct recon_params *rp, struct ct_data *data);
void rotate_slices_rotating2fixed(const struct recon_params *rp, struct ct_data *data);
void clean_up(struct recon_params *rp, struct ct_data *data);
