/*
 * openstc/mech.h -- species, thermodynamic, transport and reaction tables of
 * a finite-rate mixture, read from the reference's CHEMKIN-format inputs
 * (mechanism file, therm.dat, tran.dat).  Part of the C ABI of
 * libostc_hip.so; plain C, fixed-size arrays, no pointers, so one struct can
 * be copied to device memory as it is.
 *
 * The tables are the ones the reference builds on the host and copies into
 * DParameter (src/ChemData.h:8-85, src/ChemData.cpp:11-948,
 * src/DParameter.cu:66-154), with the reference's indexing: species in the
 * order of the mechanism's SPECIES list, reactions in file order after
 * duplicate pairs are merged, coefficient index 0..6 of the NASA-7 fits.
 */
#ifndef OPENSTC_MECH_H
#define OPENSTC_MECH_H
#ifdef __cplusplus
extern "C" {
#endif

#define OSTC_MAX_SPEC 16
#define OSTC_MAX_REAC 32
#define OSTC_MAX_ELEM 8

/* reaction labels (src/ChemData.h:63-71) */
#define OSTC_REAC_IRREVERSIBLE 0
#define OSTC_REAC_REVERSIBLE 1
#define OSTC_REAC_DUPLICATE 3
#define OSTC_REAC_THIRD_BODY 4
#define OSTC_REAC_LINDEMANN 5
#define OSTC_REAC_TROE3 6
#define OSTC_REAC_TROE4 7

typedef struct {
  int n_spec, n_reac, n_elem;
  char spec_name[OSTC_MAX_SPEC][16];
  char elem_name[OSTC_MAX_ELEM][4];
  int elem_comp[OSTC_MAX_SPEC][OSTC_MAX_ELEM];
  double mw[OSTC_MAX_SPEC];                     /* kg/kmol, from element weights (src/Element.cpp) */
  /* NASA-7, two ranges (Combustion2Part, src/Define.h:16; src/ChemData.cpp:224-542) */
  double t_low[OSTC_MAX_SPEC], t_mid[OSTC_MAX_SPEC], t_high[OSTC_MAX_SPEC];
  double low_coeff[OSTC_MAX_SPEC][7], high_coeff[OSTC_MAX_SPEC][7];
  /* transport (src/ChemData.cpp:544-635) */
  double lj_potent_inv[OSTC_MAX_SPEC];          /* k_B / epsilon */
  double vis_coeff[OSTC_MAX_SPEC];              /* 2.6693e-6 sqrt(W) / sigma^2 */
  double wj_div_wi_to_one4th[OSTC_MAX_SPEC][OSTC_MAX_SPEC];
  double sqrt_wi_div_wj_pl1_mul8[OSTC_MAX_SPEC][OSTC_MAX_SPEC];
  double binary_diffusivity_coeff[OSTC_MAX_SPEC][OSTC_MAX_SPEC];
  double kb_over_eps_jk[OSTC_MAX_SPEC][OSTC_MAX_SPEC];
  /* reactions (src/ChemData.cpp:670-948); Ea and Ea2 already divided by R_c */
  int label[OSTC_MAX_REAC];
  int rev_type[OSTC_MAX_REAC];
  int order[OSTC_MAX_REAC];                     /* sum(nu'') - sum(nu') */
  int stoi_f[OSTC_MAX_REAC][OSTC_MAX_SPEC];
  int stoi_b[OSTC_MAX_REAC][OSTC_MAX_SPEC];
  double A[OSTC_MAX_REAC], b[OSTC_MAX_REAC], Ea[OSTC_MAX_REAC];
  double A2[OSTC_MAX_REAC], b2[OSTC_MAX_REAC], Ea2[OSTC_MAX_REAC];
  double third_body_coeff[OSTC_MAX_REAC][OSTC_MAX_SPEC];
  double troe_alpha[OSTC_MAX_REAC], troe_t3[OSTC_MAX_REAC], troe_t1[OSTC_MAX_REAC], troe_t2[OSTC_MAX_REAC];
} ostc_mech_tables;

/* Reads the three files the reference reads for `species = 1`
 * (Species::Species, src/ChemData.cpp:11-106; Reaction::Reaction,
 * :670-767).  reaction = 1 parses the REACTIONS section, 0 leaves
 * n_reac = 0.  Returns 0, or a negative status with a message in err.
 * Deviations from the reference parser: none on well-formed files; the
 * reference's behaviour on a second therm.dat entry for an already-read
 * species (it also consumes the following line) is reproduced. */
int ostc_mech_load(const char* mechanism_file, const char* therm_file, const char* transport_file, int reaction,
                   ostc_mech_tables* out, char* err, int err_cap);

#ifdef __cplusplus
}
#endif
#endif
