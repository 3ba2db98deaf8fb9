function [Uhat] = Subpb2(DICmesh,beta,mu,U,F,udual,vdual,alpha,GaussPtOrder)
% Library-backed replacement of Subpb2.m:1-197 (called main_ALDIC.m:287,304,393): the Q4
% global step, assembled and solved on the GPU (fast diagonalisation on the uniform mesh
% of MeshSetUp, FP64 PCG otherwise).  GaussPtOrder is ignored, as in the reference.
aldic_mex('set_mesh', DICmesh.coordinatesFEM, DICmesh.elementsFEM(:, 1:4));
[Uhat, ~] = aldic_mex('subpb2', beta, mu, alpha, U(:), F(:), udual(:), vdual(:), 1e-12, 20000);
end
