/*
 * crabsaxs_tables.h -- published physical constant tables used on the
 * profile path (pure data, C89-compatible, no code).
 *
 * Provenance of every number (the values are those of the published sources
 * the reference also transcribes, so results agree bit for bit):
 *   - Cromer-Mann 4-Gaussian coefficients: International Tables for
 *     Crystallography Vol. C, Table 6.1.1.4 (DABAX f0_InterTables.dat);
 *     same values as crabSAXS src/formfactor.rs:52-121.
 *   - United-atom 5-Gaussian coefficients and implicit hydrogen counts:
 *     Grudinin et al. (2017) Acta Cryst. D73, Table 2;
 *     same values as crabSAXS src/formfactor.rs:185-278.
 *   - Displaced (Fraser) volumes: src/solvent.rs:406-424.
 *   - Van der Waals radii (Bondi + ionic extensions): src/solvent.rs:235-253.
 *
 * Index conventions:
 *   element  : CSX_EL_* below, same order as crabSAXS `Element`
 *              (src/structure.rs:90-106).
 *   group    : CSX_UA_* below, same order as crabSAXS `UnitedAtomGroup`
 *              (src/formfactor.rs:167-183); -1 means "no united group".
 */
#ifndef CRABSAXS_TABLES_H
#define CRABSAXS_TABLES_H

enum {
    CSX_EL_H = 0, CSX_EL_C, CSX_EL_N, CSX_EL_O, CSX_EL_S, CSX_EL_P,
    CSX_EL_NA, CSX_EL_MG, CSX_EL_CL, CSX_EL_CA, CSX_EL_FE, CSX_EL_ZN,
    CSX_EL_K, CSX_EL_I, CSX_EL_OTHER,
    CSX_NUM_ELEMENTS
};

enum {
    CSX_UA_CSP3H = 0, CSX_UA_CSP3H2, CSX_UA_CSP3H3, CSX_UA_CAROMATICH,
    CSX_UA_OALCOHOLH, CSX_UA_OACIDH, CSX_UA_ORESONANCE, CSX_UA_NH,
    CSX_UA_NH2, CSX_UA_NHPOSITIVE, CSX_UA_NH3POSITIVE, CSX_UA_GUANIDINENH,
    CSX_UA_GUANIDINENH2, CSX_UA_SH,
    CSX_NUM_UA_GROUPS
};

/* Cromer-Mann: f(s) = sum_k a[k] exp(-b[k] s^2) + c, s = q / (4 pi). */
static const double CSX_CM_A[CSX_NUM_ELEMENTS][4] = {
    {0.493002, 0.322912, 0.140191, 0.04081},   /* H  */
    {2.31, 1.02, 1.5886, 0.865},               /* C  */
    {12.2126, 3.1322, 2.0125, 1.1663},         /* N  */
    {3.0485, 2.2868, 1.5463, 0.867},           /* O  */
    {6.9053, 5.2034, 1.4379, 1.5863},          /* S  */
    {6.4345, 4.1791, 1.7800, 1.4908},          /* P  */
    {4.7626, 3.1736, 1.2674, 1.1128},          /* Na */
    {5.4204, 2.1735, 1.2269, 2.3073},          /* Mg */
    {11.4604, 7.1964, 6.2556, 1.6455},         /* Cl */
    {8.6266, 7.3873, 1.5899, 1.0211},          /* Ca */
    {11.7695, 7.3573, 3.5222, 2.3045},         /* Fe */
    {14.0743, 7.0318, 5.1652, 2.41},           /* Zn */
    {8.2186, 7.4398, 1.0519, 0.8659},          /* K  */
    {20.1472, 18.9949, 7.5138, 2.2735},        /* I  */
    {0.0, 0.0, 0.0, 0.0}                       /* Other */
};
static const double CSX_CM_B[CSX_NUM_ELEMENTS][4] = {
    {10.5109, 26.1257, 3.14236, 57.7997},      /* H  */
    {20.8439, 10.2075, 0.5687, 51.6512},       /* C  */
    {0.0057, 9.8933, 28.9975, 0.5826},         /* N  */
    {13.2771, 5.7011, 0.3239, 32.9089},        /* O  */
    {1.4679, 22.2151, 0.2536, 56.172},         /* S  */
    {1.9067, 27.1570, 0.5260, 68.1645},        /* P  */
    {3.285, 8.8422, 0.3136, 129.424},          /* Na */
    {2.8275, 79.2611, 0.3808, 7.1937},         /* Mg */
    {0.0104, 1.1662, 18.5194, 47.7784},        /* Cl */
    {10.4421, 0.6599, 85.7484, 178.437},       /* Ca */
    {4.7611, 0.3072, 15.3535, 76.8805},        /* Fe */
    {3.2655, 0.2333, 10.3163, 58.7097},        /* Zn */
    {12.7949, 0.7748, 213.187, 41.6841},       /* K  */
    {4.347, 0.3814, 27.766, 66.8776},          /* I  */
    {0.0, 0.0, 0.0, 0.0}                       /* Other */
};
static const double CSX_CM_C[CSX_NUM_ELEMENTS] = {
    0.003038, 0.2156, -11.529, 0.2508, 0.8669, 1.1149, 0.676, 0.8584,
    -9.5574, 1.3751, 1.0369, 1.3041, 1.4228, 4.0712, 0.0
};

/* United-atom groups: f(s) = sum_k a[k] exp(-b[k] s^2) + c (5 Gaussians). */
static const double CSX_UA_A[CSX_NUM_UA_GROUPS][5] = {
    {2.909530, 0.485267, 1.516151, 0.206905, 1.541626},   /* CSp3H        */
    {3.275723, 0.870037, 1.534606, 0.395078, 1.544562},   /* CSp3H2       */
    {3.681341, 1.228691, 1.549320, 0.574033, 1.554377},   /* CSp3H3       */
    {2.168070, 1.275811, 1.561096, 0.742395, -6.151144},  /* CAromaticH   */
    {0.456221, 3.219608, 0.812773, 2.666928, 1.380927},   /* OAlcoholH    */
    {3.213280, 0.463019, 0.815724, 2.664450, 1.384266},   /* OAcidH       */
    {0.688944, 2.929687, 0.416472, 2.606983, 1.319232},   /* OResonance   */
    {1.650531, 0.429639, 2.144736, 1.851894, 1.408921},   /* NH           */
    {1.904157, 1.942536, 2.435585, 0.730512, 1.379728},   /* NH2          */
    {1.426540, 0.426903, 1.878894, 1.608251, 1.200216},   /* NHPositive   */
    {1.882162, 1.933200, 2.465843, 0.927311, 1.190889},   /* NH3Positive  */
    {3.630164, 0.228310, 1.869734, 0.170550, 1.440894},   /* GuanidineNH  */
    {1.792216, 0.724464, 2.347044, 1.903020, 1.313042},   /* GuanidineNH2 */
    {0.570042, 6.337416, 1.641643, 5.398549, 1.527982}    /* SH           */
};
static const double CSX_UA_B[CSX_NUM_UA_GROUPS][5] = {
    {13.933084, 23.221524, 41.990403, 4.974183, 0.679266},
    {13.408502, 23.785175, 41.922444, 5.019072, 0.724439},
    {13.026207, 24.131974, 41.869426, 4.984373, 0.765769},
    {12.642907, 18.420069, 41.768517, 1.535360, -0.045937},
    {21.503498, 13.397134, 34.547137, 5.826620, 0.412902},
    {13.383078, 21.362223, 34.531415, 5.823549, 0.410805},
    {29.319200, 6.572228, 64.951658, 16.267799, 0.455640},
    {10.603730, 6.987283, 29.939901, 10.573859, 0.611678},
    {10.803702, 10.792421, 29.610479, 6.847755, 0.709687},
    {10.652268, 7.017651, 29.878525, 10.619493, 0.631765},
    {10.975157, 10.956008, 29.208572, 6.663555, 0.843650},
    {10.267139, 25.118086, 30.241288, 3.412776, 0.486644},
    {10.830060, 6.846763, 29.579607, 10.800018, 0.720448},
    {11.447986, 1.197657, 55.401032, 22.420955, 2.356552}
};
static const double CSX_UA_C[CSX_NUM_UA_GROUPS] = {
    0.337670, 0.377096, 0.409294, 7.400917, 0.463202, 0.458919, 0.537548,
    0.510589, 0.603738, 0.456024, 0.597322, 0.323504, 0.583312, 1.523944
};
/* Covalently attached hydrogens represented by each united group. */
static const double CSX_UA_HYDROGENS[CSX_NUM_UA_GROUPS] = {
    1.0, 2.0, 3.0, 1.0, 1.0, 1.0, 0.0, 1.0, 2.0, 1.0, 3.0, 1.0, 2.0, 1.0
};

/* Fraser displaced volumes, A^3. */
static const double CSX_EXCLUDED_VOLUME[CSX_NUM_ELEMENTS] = {
    5.15, 16.44, 2.49, 9.13, 26.09, 3.37, 4.45, 1.56, 24.84, 4.19, 7.99,
    9.85, 11.01, 44.60, 0.0
};

/* Van der Waals radii, A. */
static const double CSX_VDW_RADIUS[CSX_NUM_ELEMENTS] = {
    1.20, 1.70, 1.55, 1.52, 1.80, 1.80, 2.27, 1.73, 1.75, 2.31, 1.72,
    1.39, 2.75, 1.98, 1.70
};

/* Mean united-atom radius of the global Fraser scaling (solvent.rs:15). */
#define CSX_MEAN_ATOMIC_RADIUS 1.62
/* O-H distance of the orientationally averaged water amplitude
 * (solvent.rs:227-231). */
#define CSX_WATER_OH_DISTANCE 0.9572

#endif /* CRABSAXS_TABLES_H */
