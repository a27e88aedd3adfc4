# dataset tiny_gset
# instance g7.txt
# run_name SpecAnn
# seed 7
# n_original 7
# n_solve 7
# id_base 1
# bias_node -1
# gauge_flipped 0
# weights_negated 1
# J_sum -18.0
# energy 1.0
# cut_value 4.0
# columns node_id spin
1 -1
2 +1
3 -1
4 -1
5 +1
6 -1
7 -1
