function [StrainNodalPt] = funGlobal_NodalStrainAvg(coordinatesFEM,elementsFEM,U,GaussPtOrder)
% Library-backed replacement of funGlobal_NodalStrainAvg.m:1-147: 2x2 Gauss-point
% gradients extrapolated to the nodes and averaged, written in the reference's slot order
% [u_x u_y v_x v_y] (the off-diagonals of F's [F11 F21 F12 F22] come back swapped).
aldic_mex('set_mesh', coordinatesFEM, elementsFEM(:, 1:4));
StrainNodalPt = aldic_mex('nodal_strain', U(:), 0);
end
