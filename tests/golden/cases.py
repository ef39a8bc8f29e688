"""Golden frame cases: (name, preset, seed, synth overrides[, frame-parameter
overrides applied to the generated syntax]). Presets are the
workloads of include/hevcr_synth.h; overrides steer the generator into corners
(band positions near 31, deblocking / chroma QP offsets, intra CUs in inter
frames, 10-bit, CTB 16/32, picture sizes that leave partial CTBs, scaling
lists)."""

CASES = [
    ("intra416_s1", 0, 1, {}),
    ("intra416_s2", 0, 2, {}),
    ("intra416_s3", 0, 3, {}),
    ("intra416_s4", 0, 4, {}),
    ("intra416_nosmooth", 0, 5, {"strong_intra_smoothing": 0}),
    ("intra416_saooff", 0, 6, {"p_sao_off": 0.4}),
    ("intra416_deep_split", 0, 7, {"p_cu_split": 0.9, "p_tu_split": 0.8}),
    ("intra416_no_split", 0, 8, {"p_cu_split": 0.1, "p_tu_split": 0.1}),
    ("intra200x136_ctb32", 0, 9, {"width": 200, "height": 136, "ctb_log2": 5}),
    ("intra200x136_ctb16", 0, 10, {"width": 200, "height": 136, "ctb_log2": 4}),
    ("intra_qp_low", 0, 11, {"qp_min": 0, "qp_max": 12}),
    ("intra_qp_high", 0, 12, {"qp_min": 45, "qp_max": 51}),
    ("intra416_10bit", 0, 13, {"bit_depth": 10, "level_mul": 4, "qp_min": -12, "qp_max": 51}),
    ("intra_nodeblock", 0, 14, {"deblocking_disabled": 1}),
    ("intra_filter_offsets", 0, 15, {}, {"beta_offset_div2": 3, "tc_offset_div2": -2, "pps_cb_qp_offset": 5,
                                        "pps_cr_qp_offset": -4, "slice_cb_qp_offset": -3, "slice_cr_qp_offset": 6}),
    ("intra_filter_offsets2", 0, 16, {"qp_min": 30, "qp_max": 51}, {"beta_offset_div2": -6, "tc_offset_div2": 6,
                                                                   "pps_cb_qp_offset": 12, "pps_cr_qp_offset": -12}),
    ("b1080_s1", 1, 1, {}),
    ("b1080_s2", 1, 2, {}),
    ("p1080_s1", 5, 1, {}),
    ("b1080_intra_mix", 1, 3, {"p_intra_cu": 0.25}),
    ("b1080_unweighted", 1, 4, {"weighted": 0}),
    ("b1080_offsets", 1, 8, {"p_intra_cu": 0.1}, {"beta_offset_div2": 2, "tc_offset_div2": 4, "pps_cb_qp_offset": -7,
                                                  "slice_cr_qp_offset": 9}),
    ("b640x360_far_mv", 1, 5, {"width": 640, "height": 360, "mv_sigma_qpel": 400, "mv_clamp_qpel": 2048}),
    ("b640x360_10bit", 1, 6, {"width": 640, "height": 360, "bit_depth": 10, "level_mul": 4, "qp_min": 17}),
    ("p640x360_all_skip", 5, 7, {"width": 640, "height": 360, "p_skip": 0.95}),
    ("intra416_scaling", 0, 17, {"scaling_lists": 1}),
    ("intra416_scaling_10bit", 0, 18, {"scaling_lists": 1, "bit_depth": 10, "level_mul": 4, "qp_min": -12,
                                       "qp_max": 51}),
    ("b640x360_scaling", 1, 9, {"width": 640, "height": 360, "scaling_lists": 1, "p_intra_cu": 0.2}),
    ("intra200x136_ctb16_10bit", 0, 19, {"width": 200, "height": 136, "ctb_log2": 4, "bit_depth": 10, "level_mul": 4,
                                         "qp_min": -12, "qp_max": 51}),
    ("b200x136_ctb16_10bit", 1, 10, {"width": 200, "height": 136, "ctb_log2": 4, "bit_depth": 10, "level_mul": 4,
                                     "qp_min": 17, "p_intra_cu": 0.3}),
    ("intra1080_s1", 2, 1, {}),
    ("b2160_10bit_s1", 3, 1, {}),
    ("intra2160_10bit_s1", 4, 1, {}),
]
