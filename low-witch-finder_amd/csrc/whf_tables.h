// whf_tables.h — game data the phase-1 Seed Check needs, and the sampler-set layouts of the HIP path.
//
// The noise parameters are the game's NormalNoise settings: identifier, first octave, amplitudes. They are
// data and must equal the reference's table (engine/tools/lysh/src/noise_table.h:7-67, :77, :81, :85 for the
// biome stage's vegetation and offset noises, :86-114 for the terrain stage's); the identifier
// strings are hashed with MD5 in the seed derivation and must be spelled exactly.
#pragma once

namespace whf {

struct NoiseParam {
    const char *name;
    int first_octave;
    int n;
    double amps[16];
};

enum NoiseId {
    N_TEMPERATURE = 0, N_CONTINENTALNESS, N_EROSION, N_TEMPERATURE_LARGE, N_CONTINENTALNESS_LARGE,
    N_EROSION_LARGE, N_RIDGE, N_FLOODEDNESS, N_CAVE_ENTRANCE, N_SPAG3D_RARITY, N_SPAG3D_THICKNESS,
    N_SPAG3D_1, N_SPAG3D_2, N_SPAG_ROUGH_MOD, N_SPAG_ROUGH, N_CAVE_LAYER, N_CAVE_CHEESE, N_VEGETATION,
    N_VEGETATION_LARGE, N_OFFSET, N_JAGGED, N_SPAG2D, N_SPAG2D_ELEVATION, N_SPAG2D_MODULATOR, N_SPAG2D_THICKNESS,
    N_PILLAR, N_PILLAR_RARENESS, N_PILLAR_THICKNESS, N_NOODLE, N_NOODLE_THICKNESS, N_NOODLE_RIDGE_A, N_NOODLE_RIDGE_B,
    N_AQUIFER_BARRIER, N_AQUIFER_SPREAD, N_AQUIFER_LAVA, N_PARAM_COUNT
};

static const NoiseParam kNoise[N_PARAM_COUNT] = {
    {"minecraft:temperature", -10, 6, {1.5, 0.0, 1.0, 0.0, 0.0, 0.0}},
    {"minecraft:continentalness", -9, 9, {1.0, 1.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0, 1.0}},
    {"minecraft:erosion", -9, 5, {1.0, 1.0, 0.0, 1.0, 1.0}},
    {"minecraft:temperature_large", -12, 6, {1.5, 0.0, 1.0, 0.0, 0.0, 0.0}},
    {"minecraft:continentalness_large", -11, 9, {1.0, 1.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0, 1.0}},
    {"minecraft:erosion_large", -11, 5, {1.0, 1.0, 0.0, 1.0, 1.0}},
    {"minecraft:ridge", -7, 6, {1.0, 2.0, 1.0, 0.0, 0.0, 0.0}},
    {"minecraft:aquifer_fluid_level_floodedness", -7, 1, {1.0}},
    {"minecraft:cave_entrance", -7, 3, {0.4, 0.5, 1.0}},
    {"minecraft:spaghetti_3d_rarity", -11, 1, {1.0}},
    {"minecraft:spaghetti_3d_thickness", -8, 1, {1.0}},
    {"minecraft:spaghetti_3d_1", -7, 1, {1.0}},
    {"minecraft:spaghetti_3d_2", -7, 1, {1.0}},
    {"minecraft:spaghetti_roughness_modulator", -8, 1, {1.0}},
    {"minecraft:spaghetti_roughness", -5, 1, {1.0}},
    {"minecraft:cave_layer", -8, 1, {1.0}},
    {"minecraft:cave_cheese", -8, 9, {0.5, 1.0, 2.0, 1.0, 2.0, 1.0, 0.0, 2.0, 0.0}},
    {"minecraft:vegetation", -8, 6, {1.0, 1.0, 0.0, 0.0, 0.0, 0.0}},
    {"minecraft:vegetation_large", -10, 6, {1.0, 1.0, 0.0, 0.0, 0.0, 0.0}},
    {"minecraft:offset", -3, 4, {1.0, 1.0, 1.0, 0.0}},
    {"minecraft:jagged", -16, 16, {1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0}},
    {"minecraft:spaghetti_2d", -7, 1, {1.0}},
    {"minecraft:spaghetti_2d_elevation", -8, 1, {1.0}},
    {"minecraft:spaghetti_2d_modulator", -11, 1, {1.0}},
    {"minecraft:spaghetti_2d_thickness", -11, 1, {1.0}},
    {"minecraft:pillar", -7, 2, {1.0, 1.0}},
    {"minecraft:pillar_rareness", -8, 1, {1.0}},
    {"minecraft:pillar_thickness", -8, 1, {1.0}},
    {"minecraft:noodle", -8, 1, {1.0}},
    {"minecraft:noodle_thickness", -8, 1, {1.0}},
    {"minecraft:noodle_ridge_a", -7, 1, {1.0}},
    {"minecraft:noodle_ridge_b", -7, 1, {1.0}},
    {"minecraft:aquifer_barrier", -3, 1, {1.0}},
    {"minecraft:aquifer_fluid_level_spread", -5, 1, {1.0}},
    {"minecraft:aquifer_lava", -1, 1, {1.0}},
};

// Sampler sets. A set is a list of (noise, sub-sampler, octave slot) Perlin samplers built together for a
// seed; their tables sit contiguously per seed slot in HBM.
//   set E (8):  erosion, in the tiered-gate order A0 B0 A1 B1 A3 B3 A4 B4 (noise.c:387-430)
//   set C (28): temperature, ridge, cave_entrance, the six spaghetti noises  (climate gates + Entrance)
//   set L (36): cave_layer, cave_cheese, continentalness, floodedness       (cave ladder tail + last gates)
enum SetId { SET_E = 0, SET_C = 1, SET_L = 2, SET_COUNT = 3 };
constexpr int kSetSize[SET_COUNT] = {8, 28, 36};

// The noises of the gates, in the order of the NoiseDesc table handed to the kernels.
enum GateNoise {
    G_EROSION = 0, G_TEMPERATURE, G_RIDGE, G_CAVE_ENTRANCE, G_SPAG_RARITY, G_SPAG_THICKNESS, G_SPAG_1, G_SPAG_2,
    G_ROUGH_MOD, G_ROUGH, G_CAVE_LAYER, G_CAVE_CHEESE, G_CONTINENTALNESS, G_FLOODEDNESS, G_COUNT
};

}  // namespace whf
