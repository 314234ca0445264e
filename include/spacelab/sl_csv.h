/* spacelab_amd -- text files of a run in the reference's CSV layout.
 *
 * The reference writes three files per run (ball_group.hpp:421-516 sim_init_write,
 * ColliderSingleCore.cpp:258-286 and :333-370 sim_one_step/sim_looper) and reads
 * them back to continue one (ball_group.hpp:979-1077 parseSimData, loadConsts,
 * getLastLine):
 *   <prefix>constants.csv  one line "R,m,moi\n" per ball
 *   <prefix>simData.csv    header "x0,y0,z0,wx0,wy0,wz0,wmag0,vx0,vy0,vz0,bound0,x1,..."
 *                          then one row per write step, each preceded by '\n'
 *                          (no newline at the end of the file); per ball
 *                          x,y,z,wx,wy,wz,|w|,vx,vy,vz,bound with bound always 0
 *   <prefix>energy.csv     header "Time,PE,KE,E,p,L", then rows "t,PE,KE,PE+KE,|p|,|L|",
 *                          each preceded by '\n'
 * Numbers are printed as C++ streams print a double in the default float field,
 * i.e. printf "%.<digits>g": 6 significant digits everywhere except the energy rows
 * written by the step loop, which use 12 (ColliderSingleCore.cpp:42; the first
 * energy row, written by sim_init_write, has 6).  |w|, |p| and |L| are
 * sqrt(x*x + y*y + z*z) evaluated left to right (vec3.hpp:126).
 *
 * A continued run reads <prefix>constants.csv, or <prefix>Constants.csv where only that
 * exists (the Generator's files, Generator.cpp:168-170, whose energy layout
 * SL_CSV_ENERGY_HEADER_BOUND adds Bound,Unbound,m).
 *
 * Formatting functions follow snprintf: they return the length of the full text
 * (without the terminating NUL) and write at most cap bytes including the NUL, so a
 * call with buf = NULL, cap = 0 sizes the buffer.  Host code only: no GPU is used. */
#ifndef SPACELAB_SL_CSV_H
#define SPACELAB_SL_CSV_H

#ifdef __cplusplus
extern "C" {
#endif

#define SL_CSV_PROPERTIES 11   /* columns per ball in simData.csv (dust_const.hpp:45) */
#define SL_CSV_ENERGY_HEADER "Time,PE,KE,E,p,L"
/* the Generator's energy layout (Generator.cpp:195), read but never written */
#define SL_CSV_ENERGY_HEADER_BOUND "Time,PE,KE,E,p,L,Bound,Unbound,m"
#define SL_CSV_STATE_DIGITS 6
#define SL_CSV_ENERGY_DIGITS 12

/* header line of simData.csv for n balls (ball_group.hpp:461-472) */
long long sl_csv_sim_header(int n, char* buf, long long cap);
/* one simData.csv row, without the leading '\n' (ColliderSingleCore.cpp:268-281) */
long long sl_csv_sim_row(int n, const double* pos3, const double* w3, const double* vel3, char* buf,
                         long long cap);
/* one energy.csv row, without the leading '\n' (ColliderSingleCore.cpp:335-338);
 * p3 and L3 are the momentum and angular-momentum vectors */
long long sl_csv_energy_row(double time, double PE, double KE, const double* p3, const double* L3, int digits,
                            char* buf, long long cap);
/* the whole constants.csv text (ball_group.hpp:474-477) */
long long sl_csv_constants(int n, const double* R, const double* m, const double* moi, char* buf, long long cap);

/* Number of balls in a simData.csv row: commas / SL_CSV_PROPERTIES + 1 (ball_group.hpp:984). */
int sl_csv_count_balls(const char* row);
/* Positions, angular velocities and velocities of n balls from a simData.csv row
 * (parseSimData, ball_group.hpp:980-1011); SL_E_ARG if a field is missing or not a number. */
int sl_csv_parse_sim_row(const char* row, int n, double* pos3, double* w3, double* vel3);
/* R, m, moi of n balls from constants.csv text (loadConsts, ball_group.hpp:1018-1040). */
int sl_csv_parse_constants(const char* text, int n, double* R, double* m, double* moi);
/* Columns of an energy.csv file from its header line: 6 for SL_CSV_ENERGY_HEADER, 9 for
 * SL_CSV_ENERGY_HEADER_BOUND (the Generator's, Generator.cpp:195); SL_E_ARG otherwise. */
int sl_csv_energy_columns(const char* header);
/* The ncols (6 or 9) numbers of one energy.csv row, which must hold exactly that many. */
int sl_csv_parse_energy_row(const char* row, int ncols, double* out);
/* The last line of a file as getLastLine reads it (ball_group.hpp:1043-1077): the text
 * after the final '\n' (empty if the file ends with one, or if that '\n' is its first
 * byte), or the whole file if it has none.  Returns its length (snprintf rules), SL_E_ARG if the file cannot be read. */
long long sl_csv_last_line(const char* path, char* buf, long long cap);

#ifdef __cplusplus
}
#endif

#endif
