Elements outside the batched NR path
R1 1 0 1k
L1 1 2 1u
K1 L1 L1 0.5
E1 3 0 1 0 2
G1 3 0 1 0 1m
Q1 4 1 0 npn1
.subckt buf a y
R1 a y 1k
.ends
X1 1 5 buf
.end
