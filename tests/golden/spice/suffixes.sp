Value suffixes and continuation
R1 1 2 1MEG
R2 2 3 2.2meg
R3 3 4 10mil
R4 4 5 3.3k
R5 5 6 4u7
R6 6 0 47
+ extra-token
C1 1 0 1.5f
C2 2 0 2e-12
C3 3 0 100n
C4 4 0 1p
V1 1 0 1.8V
I2 0 5 5a
R7 6 7 1t
R8 7 0 2g
.END
