// internal.h -- declarations shared by the library's translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "bubblesim.h"

namespace bs {

// thread-local last error (capi.cpp); returns rc for chaining
int set_error(int rc, const std::string &msg);
void clear_error();

// host physics (physics.cpp)
void model_default(bs_model *m);
float model_eps(const bs_model &m);
float model_m2(const bs_model &m);
float model_delta(const bs_model &m);
float model_wall_width(const bs_model &m);
float model_surface_tension(const bs_model &m);
float model_critical_radius(const bs_model &m);
float model_e_ref(const bs_model &m);
float model_t_ref(const bs_model &m);
float model_a_rad(const bs_model &m);
float model_dt(const bs_model &m);
float model_seed_radius(const bs_model &m);
int model_validate(const bs_model &m, std::string &err);
void model_derive(const bs_model &m, bs_derived *d);
float accumulated_time(uint64_t k, float dt);
int make_schedule(int32_t mode, uint32_t count, float radius, const uint32_t grid[3], float duration, uint64_t seed,
                  uint32_t uniform_max_step, uint64_t time_seed, float dt, std::vector<bs_event> &out,
                  std::string &err);
int spec_schedule(const bs_spec &s, std::vector<bs_event> &out, std::string &err);
uint64_t event_step(float t, float dt);
void sim_params_fill(const bs_model &m, const uint32_t grid[3], float time, bs_sim_params *p);
int spec_validate(const bs_spec &s, std::string &err);
void spec_default(bs_spec *s);
int parse_grid(const char *s, uint32_t g[3], std::string &err);
uint64_t bytes_per_cell(int32_t precision, int32_t kernel);

}  // namespace bs
