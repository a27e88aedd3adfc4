# dataset barabasi_albert_3ins_m10_tiny
# instance ba_20_1_0.txt
# run_name SpecAnn
# seed 7
# n_original 20
# n_solve 21
# id_base 1
# bias_node 20
# gauge_flipped 1
# weights_negated 0
# J_sum -972.0
# energy 1076.0
# cut_value 
# columns node_id spin
1 +1
2 -1
3 +1
4 +1
5 -1
6 +1
7 +1
8 -1
9 +1
10 +1
11 -1
12 +1
13 +1
14 -1
15 +1
16 +1
17 -1
18 +1
19 +1
20 -1
