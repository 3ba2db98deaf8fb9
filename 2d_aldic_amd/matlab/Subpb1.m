function [U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1(UOld,FOld,udual,vdual,coordinatesFEM,...
    Df,ImgRef,ImgDef,mu,beta,HPar,ALSolveStep,DICpara,ICGNmethod,tol)
% Library-backed replacement of Subpb1.m:1-199 (called main_ALDIC.m:361): 2-DOF IC-GN with
% F frozen (funICGN_Subpb1.m) on the GPU from LocalICGN's Hessian (a zero Hessian at
% ALSolveStep 1, as the reference builds it), failed nodes (status < 0 or > 99) dropped,
% then u and v refitted at every node by gridfit ('springs', the reference's options) on
% the GPU.  beta and udual are unused, as in the reference.
if nargin < 15, tol = 1e-2; end
S = size(coordinatesFEM, 1);
if iscell(HPar), H = zeros(21, S); for k = 1:21, H(k, :) = HPar{k}(:)'; end
else, H = HPar'; end
if ALSolveStep <= 1, H = zeros(21, S); end
aldic_mex('set_normalized', double(ImgRef), double(ImgDef));
aldic_mex('set_gradients', Df.DfDx, Df.DfDy);
t0 = tic;
[U, it] = aldic_mex('icgn2', coordinatesFEM, DICpara.winsize, H, mu, vdual(:), UOld(:), FOld(:), tol);
ALSub1Time = toc(t0);
ConvItPerEle = double(it);
bad = ConvItPerEle < 0 | ConvItPerEle > 99;
LocalICGNBadPtNum = nnz(bad);
U([2*find(bad)-1; 2*find(bad)]) = NaN;
xn = unique(coordinatesFEM(:, 1)); yn = unique(coordinatesFEM(:, 2));
[~, ix] = ismember(coordinatesFEM(:, 1), xn); [~, iy] = ismember(coordinatesFEM(:, 2), yn);
for c = 1:2
    Z = nan(numel(yn), numel(xn));
    Z(sub2ind(size(Z), iy, ix)) = U(c:2:end);
    Zg = aldic_mex('gridfit', xn(:)', yn(:)', Z, 1e-12, 4000);
    U(c:2:end) = Zg(sub2ind(size(Zg), iy, ix));
end
end
