Inverter pair with diode clamp
.model nch NMOS (KP=110u VTO=0.7 LAMBDA=0.04 LEVEL=1)
.model pch PMOS (KP=50u VTO=-0.7 LAMBDA=0.05)
.MODEL dmod D (IS=1e-14 N=1.05)
VDD vdd 0 1.8
VIN in 0 PULSE(0 1.8 20n 1n 1n 40n 0)
M1 out in 0 0 nch W=2u L=1u
M2 out in vdd vdd pch W=4u L=1u
M3 out2 out 0 0 nch
M4 out2 out vdd vdd pch w=8u l=2u
D1 out2 clamp dmod
D2 clamp 0 DMOD AREA=2
R1 clamp 0 10k
C1 out 0 10f
C2 out2 0 10f
M5 x in vss 0 weird_pmos_model W=1u L=1u
R2 x vdd 100k
.tran 1n 100n
.end
