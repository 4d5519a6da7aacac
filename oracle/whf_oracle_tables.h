/* whf_oracle_tables.h — the game's noise parameters used by the phase-1 Seed Check and the biome stage
 * (oracle copy).
 *
 * TEST INFRASTRUCTURE: part of the CPU oracle; never linked into the HIP product path.
 *
 * Each entry is Minecraft's NormalNoise parameter set for one identifier: the first octave and the
 * amplitude list. They are game data, not code, so the values must be identical to the reference's
 * table (engine/tools/lysh/src/noise_table.h:7-67, rows "minecraft:temperature" .. "minecraft:cave_cheese", and
 * "minecraft:offset" at noise_table.h:85 with its amplitudes LYSH_AMPS_9 at :16, and the terrain stage's noises at
 * :86-114 (jagged, spaghetti_2d*, pillar*, noodle*, aquifer_barrier/_fluid_level_spread/_lava);
 * the identifier strings feed the MD5-keyed seed derivation and must be spelled exactly).
 */
#ifndef WHF_ORACLE_TABLES_H
#define WHF_ORACLE_TABLES_H

typedef struct {
    const char *name;
    int first_octave;
    int n;
    double amps[16];
} wo_noise_param;

enum {
    WO_N_TEMPERATURE = 0,
    WO_N_VEGETATION,
    WO_N_CONTINENTALNESS,
    WO_N_EROSION,
    WO_N_TEMPERATURE_LARGE,
    WO_N_VEGETATION_LARGE,
    WO_N_CONTINENTALNESS_LARGE,
    WO_N_EROSION_LARGE,
    WO_N_RIDGE,
    WO_N_FLOODEDNESS,
    WO_N_CAVE_ENTRANCE,
    WO_N_SPAG3D_RARITY,
    WO_N_SPAG3D_THICKNESS,
    WO_N_SPAG3D_1,
    WO_N_SPAG3D_2,
    WO_N_SPAG_ROUGH_MOD,
    WO_N_SPAG_ROUGH,
    WO_N_CAVE_LAYER,
    WO_N_CAVE_CHEESE,
    WO_N_OFFSET,
    WO_N_JAGGED,
    WO_N_SPAG2D,
    WO_N_SPAG2D_ELEVATION,
    WO_N_SPAG2D_MODULATOR,
    WO_N_SPAG2D_THICKNESS,
    WO_N_PILLAR,
    WO_N_PILLAR_RARENESS,
    WO_N_PILLAR_THICKNESS,
    WO_N_NOODLE,
    WO_N_NOODLE_THICKNESS,
    WO_N_NOODLE_RIDGE_A,
    WO_N_NOODLE_RIDGE_B,
    WO_N_AQUIFER_BARRIER,
    WO_N_AQUIFER_SPREAD,
    WO_N_AQUIFER_LAVA,
    WO_N_COUNT
};

static const wo_noise_param WO_NOISE[WO_N_COUNT] = {
    { "minecraft:temperature",                    -10, 6, { 1.5, 0.0, 1.0, 0.0, 0.0, 0.0 } },
    { "minecraft:vegetation",                      -8, 6, { 1.0, 1.0, 0.0, 0.0, 0.0, 0.0 } },
    { "minecraft:continentalness",                 -9, 9, { 1.0, 1.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0, 1.0 } },
    { "minecraft:erosion",                         -9, 5, { 1.0, 1.0, 0.0, 1.0, 1.0 } },
    { "minecraft:temperature_large",              -12, 6, { 1.5, 0.0, 1.0, 0.0, 0.0, 0.0 } },
    { "minecraft:vegetation_large",               -10, 6, { 1.0, 1.0, 0.0, 0.0, 0.0, 0.0 } },
    { "minecraft:continentalness_large",          -11, 9, { 1.0, 1.0, 2.0, 2.0, 2.0, 1.0, 1.0, 1.0, 1.0 } },
    { "minecraft:erosion_large",                  -11, 5, { 1.0, 1.0, 0.0, 1.0, 1.0 } },
    { "minecraft:ridge",                           -7, 6, { 1.0, 2.0, 1.0, 0.0, 0.0, 0.0 } },
    { "minecraft:aquifer_fluid_level_floodedness", -7, 1, { 1.0 } },
    { "minecraft:cave_entrance",                   -7, 3, { 0.4, 0.5, 1.0 } },
    { "minecraft:spaghetti_3d_rarity",            -11, 1, { 1.0 } },
    { "minecraft:spaghetti_3d_thickness",          -8, 1, { 1.0 } },
    { "minecraft:spaghetti_3d_1",                  -7, 1, { 1.0 } },
    { "minecraft:spaghetti_3d_2",                  -7, 1, { 1.0 } },
    { "minecraft:spaghetti_roughness_modulator",   -8, 1, { 1.0 } },
    { "minecraft:spaghetti_roughness",             -5, 1, { 1.0 } },
    { "minecraft:cave_layer",                      -8, 1, { 1.0 } },
    { "minecraft:cave_cheese",                     -8, 9, { 0.5, 1.0, 2.0, 1.0, 2.0, 1.0, 0.0, 2.0, 0.0 } },
    { "minecraft:offset",                          -3, 4, { 1.0, 1.0, 1.0, 0.0 } },
    { "minecraft:jagged",                         -16, 16, { 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0,
                                                             1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0 } },
    { "minecraft:spaghetti_2d",                    -7, 1, { 1.0 } },
    { "minecraft:spaghetti_2d_elevation",          -8, 1, { 1.0 } },
    { "minecraft:spaghetti_2d_modulator",         -11, 1, { 1.0 } },
    { "minecraft:spaghetti_2d_thickness",         -11, 1, { 1.0 } },
    { "minecraft:pillar",                          -7, 2, { 1.0, 1.0 } },
    { "minecraft:pillar_rareness",                 -8, 1, { 1.0 } },
    { "minecraft:pillar_thickness",                -8, 1, { 1.0 } },
    { "minecraft:noodle",                          -8, 1, { 1.0 } },
    { "minecraft:noodle_thickness",                -8, 1, { 1.0 } },
    { "minecraft:noodle_ridge_a",                  -7, 1, { 1.0 } },
    { "minecraft:noodle_ridge_b",                  -7, 1, { 1.0 } },
    { "minecraft:aquifer_barrier",                 -3, 1, { 1.0 } },
    { "minecraft:aquifer_fluid_level_spread",      -5, 1, { 1.0 } },
    { "minecraft:aquifer_lava",                    -1, 1, { 1.0 } },
};

#endif /* WHF_ORACLE_TABLES_H */
